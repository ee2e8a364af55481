#!/usr/bin/env python
"""Headline benchmark: blocked Bunch-Kaufman LDL^T on one MI355X.

Workload (n = 16384): A = Q diag(lambda) Q^T with Q from the QR of a
seeded Gaussian matrix, |lambda| log-spaced over [1e-6, 1] with seeded
random signs (condition 1e6), b ~ N(0, 1).  A "step" is one plain FP64
factorization (bbk_dsytrf) of a resident copy of A.

Printed JSON line:
  value      GFLOP/s of the FP64 factorization, n^3/3 convention
  roofline   dominant kernel vs its MI355X roofline (HIP events, live),
             against the datasheet peak and the device's achievable peak
             (achievable_peaks: a copy kernel and MFMA loops, measured first)
  fp64 / mixed  time to backward error <= 1e-12 for A x = b:
             FP64 factor + solve, and FP32 factor + FP64 refinement
  batched    the 4096-matrix (n in [16, 256]) batch in both precisions and
             the 1024-matrix (n in [256, 2048]) batched mixed solve
  out_of_core  the same matrix factored from pinned host memory with a
             quarter of its bytes as the HBM budget, checked bitwise
             against the in-core factors
  cpu_baseline  the reference's own Factorize() (built from its sources,
             OpenMP on all host cores) on a bounded sample of the same
             matrix family
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from blockedbunchkaufmanldlt_amd import workloads  # noqa: E402

PEAK_FP64 = 78.6e12   # MI355X dense FP64 matrix, FLOP/s
PEAK_FP32 = 157.3e12  # MI355X dense FP32 matrix, FLOP/s
PEAK_HBM = 8.0e12     # bytes/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "factor_fp64_n16384_pmc_traffic.md")


def measured_traffic():
    """HBM bytes per launch per kernel, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy kernel)."""
    try:
        text = open(TRAFFIC_PROFILE).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(block)["traffic_bytes_per_launch"]
    except (OSError, IndexError, KeyError, ValueError):
        return {}


def bench_batched_small(torch, B, h, reps=5, seed=4096):
    """4096 matrices, n uniform in [16, 256], entries uniform [-1, 1]: one
    persistent launch replayed from a HIP graph, FP64 and FP32; panel
    widths / team sizes from the autotuner (tuning time reported apart)."""
    ns = workloads.c4_sizes(seed)
    flops = sum(n ** 3 / 3.0 for n in ns)
    out = {"workload": "4096 matrices, n uniform in [16, 256], entries uniform [-1, 1]",
           "flops": flops}
    for dt, name, peak in ((torch.float64, "fp64", PEAK_FP64), (torch.float32, "fp32", PEAK_FP32)):
        eb = 8 if dt == torch.float64 else 4
        t0 = time.perf_counter()
        h.batched_tune(dt, 256)
        tune_s = time.perf_counter() - t0
        src = h.gen_synth_batched(ns, seed=seed, dtype=dt)
        work = [s.clone() for s in src]
        perms = [torch.empty(n, dtype=torch.int32, device="cuda") for n in ns]
        pivs = [torch.empty(n, dtype=torch.int32, device="cuda") for n in ns]
        g = h.graph_sytrf_batched(work, perms, pivs)
        times = []
        for _ in range(reps):
            for a, b in zip(work, src):
                a.copy_(b)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(h.stream)
            g.launch()
            e1.record(h.stream)
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        g.close()
        ms = min(times)
        alg_bytes = 2.0 * sum(n * (n + 1) / 2.0 for n in ns) * eb
        out[name] = {"ms": round(ms, 3), "GFLOP/s": round(flops / ms / 1e6, 1),
                     "frac_matrix_peak": round(flops / (ms * 1e-3) / peak, 5),
                     "hbm_GB/s": round(alg_bytes / ms / 1e6, 1),
                     "frac_hbm": round(alg_bytes / (ms * 1e-3) / PEAK_HBM, 5),
                     "widths": h.batched_widths(dt), "team_sizes": h.batched_team_sizes(dt),
                     "update_tiles": h.batched_tiles(dt),
                     "tune_s": round(tune_s, 3)}
        del src, work, perms, pivs
    torch.cuda.empty_cache()
    return out


def bench_batched_mixed(torch, B, h, reps=2, seed=1024):
    """1024 matrices, n uniform in [256, 2048], condition 10^U[4,7]
    (A = Q diag(lambda) Q^T, Q from the QR of a Gaussian matrix), 4
    right-hand sides N(0, 1): binary32
    batched factorization + binary64 refinement to backward error 1e-12."""
    ns, mats, Bs = workloads.c5_batch(torch, B.padded_ld, seed)
    Xs = [torch.zeros((n, 4), dtype=torch.float64, device="cuda") for n in ns]
    torch.cuda.synchronize()
    best = None
    for _ in range(reps):
        for x in Xs:
            x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = h.sysv_mixed_batched(mats, Bs, Xs, tol=1e-12, max_iter=30)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if best is None or dt < best[0]:
            best = (dt, info)
    dt, info = best
    flops = sum(n ** 3 / 3.0 for n in ns)
    out = {"workload": "1024 matrices, n uniform in [256, 2048], A = Q diag(lambda) Q^T with "
                       "dense random orthogonal Q, cond 10^U[4,7], 4 RHS",
           "time_ms": round(dt * 1e3, 3), "factor_GFLOP/s_equiv": round(flops / dt / 1e9, 1),
           "converged": bool(info["converged"]), "max_berr": float(info["berr"].max()),
           "iters_mean": round(float(info["iters"].mean()), 3),
           "iters_max": int(info["iters"].max())}
    del mats, Bs, Xs
    torch.cuda.empty_cache()
    return out


def bench_ooc(torch, B, h, A0, n, nb=128, frac=4):
    """The same matrix factored out of core (bbk_dsytrf_ooc): A in pinned
    host memory, HBM for its copies bounded by 1/frac of the matrix's
    device bytes.  Checked bitwise against an in-core factorization.
    Algorithmic PCIe bytes: for every panel run on the host copy, the
    trailing upper triangle in and out plus the U rows; the trailing matrix
    in and out once it fits; the deferred gathers of the host-resident rows."""
    import numpy as np
    budget = 8 * n * B.padded_ld(n) // frac
    T = A0.clone()
    perm, piv = h.sytrf(T, nb)
    h.synchronize()
    ref = np.triu(T[:, :n].cpu().numpy())
    del T
    torch.cuda.empty_cache()
    M = B.PinnedMatrix(n)
    try:
        M.array[:, :] = A0[:, :n].cpu().numpy()
        t0 = time.perf_counter()
        p2, v2 = h.sytrf_ooc(M, nb, budget)
        dt = time.perf_counter() - t0
        same = (np.array_equal(np.triu(M.array), ref) and np.array_equal(p2, perm.cpu().numpy())
                and np.array_equal(v2, piv.cpu().numpy()))
    finally:
        M.close()
    w0 = 0
    while w0 < n and 8 * (n - w0) * B.padded_ld(n - w0) > budget:
        w0 += nb
    alg = sum((n - e) * (n - e + 1) / 2.0 * 16 + nb * (n - e) * 8.0
              for e in range(nb, w0 + 1, nb))
    alg += (n - w0) * (n - w0 + 1) / 2.0 * 16 + sum(nb * (n - e) * 16.0
                                                    for e in range(nb, w0 + 1, nb))
    return {"workload": f"n={n} FP64 in pinned host memory, HBM budget 1/{frac} of the matrix",
            "device_budget_bytes": budget, "time_ms": round(dt * 1e3, 1),
            "panels_on_host": w0 // nb, "pcie_alg_bytes": alg,
            "pcie_alg_GB/s": round(alg / dt / 1e9, 1), "bitwise_equal_in_core": bool(same)}


def achievable_peaks():
    """The device's achievable peaks (tools/bbk_peaks.hip, built beside the
    library): HBM copy GB/s and FP64 / FP32 MFMA TFLOP/s.  A child process,
    run before this process touches the GPU."""
    exe = os.path.join(ROOT, "blockedbunchkaufmanldlt_amd", "_build", "bbk_peaks")
    if not os.path.exists(exe):
        return None
    try:
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    except (OSError, subprocess.TimeoutExpired):
        return None
    if r.returncode != 0 or not r.stdout.strip():
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(sizes=(4096, 8192), bs=128):
    """The reference's Factorize() (oracle/_ref, OpenMP, all host cores) on a
    bounded sample: two orders, so the rate's trend with n shows (the
    headline n = 16384 would take minutes); `value` is the larger order's
    rate.  Falls back to the oracle restatement if the reference checker
    was not built."""
    import oracle as O
    cores = os.cpu_count() or 1
    rates, total = {}, 0.0
    kind, what = "reference", "reference BlockLdltDecomposition::Factorize() (kCompactOnly, OpenMP)"
    for n_sample in sizes:
        A = O.gen_synth_symmetric(n_sample, 424242, 1.0, 7)
        try:
            if not O.ref_available():
                raise FileNotFoundError
            secs = O.ref_time_factorize(A, bs)
        except Exception:
            t0 = time.perf_counter()
            O.dsytrf(A, bs, omp=True)
            secs = time.perf_counter() - t0
            kind, what = "port", "oracle restatement of Factorize() (OpenMP)"
        rates[n_sample] = n_sample ** 3 / 3.0 / secs / 1e9
        total += secs
        del A
    n_top = sizes[-1]
    trend = ", ".join(f"n={n}: {r:.2f} GFLOP/s" for n, r in rates.items())
    out = {"value": round(rates[n_top], 3), "unit": "GFLOP/s", "cores": cores, "kind": kind,
           "by_n": {str(n): round(r, 3) for n, r in rates.items()},
           "sample": f"{what}, n in {list(sizes)} (value at n={n_top}), bs={bs}, uniform[-1,1] "
                     f"entries, zero diagonal every 7th index; {trend}; {total:.1f} s on {cores} "
                     f"host threads"}
    # the 4096-matrix batch: the same matrices as the GPU's batched leg,
    # an outer OpenMP loop over the matrices, each factorization on one thread
    if kind == "reference":
        ns = workloads.c4_sizes(4096)
        mats = [O.gen_synth_symmetric(nm, 4096 + m, 1.0, 0) for m, nm in enumerate(ns)]
        secs = O.ref_time_batch(mats, 64)
        flops = sum(nm ** 3 / 3.0 for nm in ns)
        out["batched_small"] = {
            "value": round(flops / secs / 1e9, 3), "unit": "GFLOP/s", "ms": round(secs * 1e3, 1),
            "sample": "the reference's Factorize() (kCompactOnly, bs=64) on all 4096 matrices "
                      "of the small batch (the GPU leg's inputs), omp parallel for "
                      "schedule(dynamic) over the matrices, each factorization on one thread"}
        del mats
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--seed", type=int, default=20260)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-ooc", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="factorizations only (for rocprofv3 runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the factorization's pivot chain is not sharded across GPUs)", file=sys.stderr)
        return 2

    peaks = None
    if not args.profile:
        log("achievable peaks")
        peaks = achievable_peaks()

    import torch
    import blockedbunchkaufmanldlt_amd as B

    n = args.n
    lda = B.padded_ld(n)
    h = B.Handle()
    log(f"generating n={n} eigenvalue-constructed matrix")
    A0, b = workloads.c3_matrix(torch, n, args.seed, lda)
    torch.cuda.synchronize()

    # ---- timed FP64 factorizations on resident copies -------------------
    W, K = args.warmup, args.steps
    copies = [A0.clone() for _ in range(W + K)]
    torch.cuda.synchronize()
    for k in range(W):
        log(f"warmup {k}")
        h.sytrf(copies[k])
    h.synchronize()
    h.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        last = h.sytrf(copies[W + k])
    h.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    tm = h.last_timing()
    h.set_timing(False)
    log(f"timed {K} factorizations: {elapsed * 1e3 / K:.2f} ms each; kernels {tm}")
    # verification of the last timed factorization (outside the timed region):
    # ||P L D L^T P^T - A||_F / max(||A||_F, 1) and the reference stress
    # test's sampled residual, both on the device
    verification = None
    if K > 0:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res_full = h.residual(A0, copies[W + K - 1], last[0], last[1])
        res_ms = (time.perf_counter() - t1) * 1e3
        res_sampled = h.sampled_residual(A0, copies[W + K - 1], last[0], last[1], args.seed)
        log(f"residual full {res_full:.3e} ({res_ms:.1f} ms), sampled {res_sampled:.3e}")
        verification = {"relative_residual": res_full,
                        "sampled_relative_residual": res_sampled,
                        "residual_ms": round(res_ms, 2),
                        "n_2x2_pivots": int((last[1] == 2).sum())}
        del last
    del copies
    torch.cuda.empty_cache()

    flops = n ** 3 / 3.0
    ms_per_step = elapsed * 1e3 / K
    gflops = flops / (elapsed / K) / 1e9

    # roofline of the kernels (per-launch averages over the timed region)
    panels_per_step = tm["panels"] / K
    upd_ms = tm["update_ms"] / K
    pan_ms = tm["panel_ms"] / K
    upd_flops = flops  # the trailing updates carry the n^3/3 flops
    upd_launches = panels_per_step
    upd_avg_ms = upd_ms / max(upd_launches, 1)
    upd_achieved = upd_flops / (upd_ms * 1e-3) / 1e12 if upd_ms > 0 else 0.0
    # panel kernel: algorithmic bytes = read A rows k (+ candidate rows) and
    # write U rows + W rows, 8 B each: about 4 n^2 / 2 * 8 B per factorization
    pan_bytes = 4.0 * n * n / 2 * 8
    pan_achieved = pan_bytes / (pan_ms * 1e-3) / 1e9 if pan_ms > 0 else 0.0
    roof_update = {"kernel": "bk_trailing_update_kernel<double>", "bound": "mfma",
                   "achieved": round(upd_achieved, 3), "peak": PEAK_FP64 / 1e12,
                   "unit": "TFLOP/s", "frac": round(upd_achieved / (PEAK_FP64 / 1e12), 4),
                   "traffic": None, "avg_launch_ms": round(upd_avg_ms, 4),
                   "launches_per_step": upd_launches}
    roof_panel = {"kernel": "bk_panel_kernel<double>", "bound": "hbm",
                  "achieved": round(pan_achieved, 3), "peak": PEAK_HBM / 1e9,
                  "unit": "GB/s", "frac": round(pan_achieved / (PEAK_HBM / 1e9), 5),
                  "traffic": None, "avg_launch_ms": round(pan_ms / max(panels_per_step, 1), 4),
                  "us_per_pivot": round(pan_ms * 1e3 / n, 3)}
    traffic = measured_traffic() if n == 16384 else {}
    for roof, key in ((roof_update, "bk_trailing_update_kernel<double>"),
                      (roof_panel, "bk_panel_kernel<double>")):
        if key in traffic:
            roof["traffic"] = traffic[key]
            roof["traffic_source"] = os.path.relpath(TRAFFIC_PROFILE, ROOT)
    if peaks:
        # the same achieved figures against the device's measured peaks
        roof_update["peak_achievable"] = peaks["fp64_mfma_TFLOP/s"]
        roof_update["frac_achievable"] = round(upd_achieved / peaks["fp64_mfma_TFLOP/s"], 4)
        roof_panel["peak_achievable"] = peaks["hbm_copy_GB/s"]
        roof_panel["frac_achievable"] = round(pan_achieved / peaks["hbm_copy_GB/s"], 5)
    dominant = roof_update if upd_ms >= pan_ms else roof_panel

    # ---- time to backward error 1e-12: FP64 vs mixed ----------------------
    solve = {"fp64": None, "mixed": None, "speedup_mixed_vs_fp64": None}
    if not args.no_solve and not args.profile:
        X = torch.zeros_like(b)
        try:
            # two untimed runs, then the median (and the best) of five
            res = []
            for rep in range(7):
                torch.cuda.synchronize(); t0 = time.perf_counter()
                be = h.sysv(A0, b, X)
                torch.cuda.synchronize(); res.append((time.perf_counter() - t0, be))
            res = sorted(res[2:], key=lambda r: r[0])
            t64, be64 = res[2]
            t64_min = res[0][0]
            res = []
            for rep in range(7):
                X.zero_()
                torch.cuda.synchronize(); t0 = time.perf_counter()
                info = h.sysv_mixed(A0, b, X, tol=1e-12, max_iter=30)
                torch.cuda.synchronize(); res.append((time.perf_counter() - t0, info))
            res = sorted(res[2:], key=lambda r: r[0])
            tmx, infomx = res[2]
            tmx_min = res[0][0]
            solve = {"fp64": {"time_ms": round(t64 * 1e3, 3), "best_ms": round(t64_min * 1e3, 3),
                              "berr": be64},
                     "mixed": {"time_ms": round(tmx * 1e3, 3), "best_ms": round(tmx_min * 1e3, 3),
                               "berr": infomx["berr"], "iters": infomx["iters"],
                               "converged": infomx["converged"]},
                     "speedup_mixed_vs_fp64": round(t64 / tmx, 3),
                     "timing": "median of 5 after 2 untimed runs, wall clock around "
                               "synchronized calls"}
        except B.BbkError as e:
            log(f"solve path unavailable: {e}")

    batched = None
    if not args.no_batched and not args.profile:
        log("batched workloads")
        batched = {"small_fp64_fp32": bench_batched_small(torch, B, h),
                   "mixed_refined": bench_batched_mixed(torch, B, h)}

    ooc = None
    if not args.no_ooc and not args.profile:
        log("out-of-core factorization")
        ooc = bench_ooc(torch, B, h, A0, n)

    cpu = None
    if not args.no_cpu and not args.profile:
        log("cpu baseline")
        cpu = cpu_baseline()

    out = {
        "metric": "GFLOP/s (n^3/3) and time to 1e-12 backward error, mixed vs FP64, 1 GPU",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"FP64 blocked Bunch-Kaufman LDL^T, n={n}, "
                               f"A=Q diag(lambda) Q^T (cond 1e6)",
                   "n": n, "lda": lda, "panel_width": -(-n // int(panels_per_step)),
                   "parallelism": "single GPU"},
        "roofline": dominant,
        "roofline_update": roof_update,
        "roofline_panel": roof_panel,
        "achievable_peaks": peaks,
        "fraction_of_fp64_peak": round(gflops * 1e9 / PEAK_FP64, 4),
        "fraction_of_fp64_achievable": (round(gflops / 1e3 / peaks["fp64_mfma_TFLOP/s"], 4)
                                        if peaks else None),
        "time_to_1e-12": solve,
        "verification": verification,
        "batched": batched,
        "out_of_core": ooc,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    # release device state before interpreter teardown (profilers attached
    # to the process write their results after this point)
    del A0, b
    h.close()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""GPU parity tests for the batched persistent factorization
(bbk_?sytrf_batched, bbk_batched_tune, bbk_graph_*).

Every matrix of a batch is checked against the oracle at the same panel
width (the batched kernel keeps the reference's nominal panel boundaries
and defers a straddling 2x2 to the next panel exactly as the reference
does):
* pivot types and permutation identical whenever every decision is well
  separated (margin > 1e-8), compact factors to 1e-9 relative;
* residual < 1e-10 and within 10x of the oracle's (FP64); FP32 factors:
  pivots equal to the FP64 decisions on well-separated inputs, residual at
  binary32 level.
Inputs come from bbk_?gen_synth_batched, whose matrix m is the SplitMix64
formula with seed + m, restated by oracle.gen_synth_symmetric.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
MARGIN = 1e-8


def host_matrix(T):
    n = T.shape[0]
    A = T[:, :n].double().cpu().numpy()
    return np.triu(A) + np.triu(A, 1).T


def factor_batch(handle, oracle, mats, dtype):
    import torch
    perms = [torch.empty(m.shape[0], dtype=torch.int32, device="cuda") for m in mats]
    pivs = [torch.empty(m.shape[0], dtype=torch.int32, device="cuda") for m in mats]
    handle.sytrf_batched(mats, perms, pivs)
    handle.synchronize()
    out = []
    for m, p, v in zip(mats, perms, pivs):
        n = m.shape[0]
        F = np.triu(m[:, :n].double().cpu().numpy())
        out.append(oracle.Factorization(n, 0, F, p.cpu().numpy().astype(np.int32),
                                        v.cpu().numpy().astype(np.int32)))
    return out


def width_for(widths, n):
    c = 0
    while c < 7 and n > (32 << c):
        c += 1
    return widths[c]


def check(oracle, A, g, bs, fp32=False):
    o = oracle.dsytrf(A, bs, trace=True)
    res_g = oracle.relative_residual(A, g)
    res_o = oracle.relative_residual(A, o)
    if fp32:
        assert res_g < 5e-5 * max(1, A.shape[0] / 64), res_g
    else:
        assert res_g < 1e-10 and res_g <= 10 * max(res_o, 1e-15), (res_g, res_o)
    if oracle.min_margin(o.trace) > (1e-3 if fp32 else MARGIN):
        assert np.array_equal(g.piv, o.piv)
        assert np.array_equal(g.perm, o.perm)
        if not fp32:
            scale = max(1.0, np.abs(o.compact).max())
            assert np.abs(g.compact - o.compact).max() / scale < 1e-9
    return o


def test_batched_fp64_matches_oracle(oracle, handle):
    import torch
    ns = [1, 2, 3, 7, 16, 31, 32, 33, 47, 64, 65, 100, 127, 128, 129, 200, 255, 256, 257, 300,
          17, 90, 150, 220]
    mats = handle.gen_synth_batched(ns, seed=5150, scale=1.0, zero_diag_stride=7,
                                    dtype=torch.float64, pad=3)
    hosts = [host_matrix(m) for m in mats]
    for m, A in enumerate(hosts):
        assert np.array_equal(A, oracle.gen_synth_symmetric(ns[m], 5150 + m, 1.0, 7))
    widths = handle.batched_widths(torch.float64)
    outs = factor_batch(handle, oracle, mats, torch.float64)
    for A, g in zip(hosts, outs):
        check(oracle, A, g, width_for(widths, A.shape[0]))


@pytest.mark.parametrize("bw,nt", [(1, 256), (2, 64), (3, 256), (5, 64)])
def test_batched_small_widths_defer_2x2(oracle, handle, bw, nt):
    """Narrow panels make 2x2 pivots straddle panel edges often; the
    reference defers them (the next panel starts one column early).  Both
    team sizes (one wavefront / four per matrix)."""
    import torch
    saved_w = handle.batched_widths(torch.float64)
    saved_nt = handle.batched_team_sizes(torch.float64)
    handle.set_batched_widths(torch.float64, [bw] * 8)
    handle.set_batched_team_sizes(torch.float64, [nt] * 8)
    try:
        ns = [9, 16, 31, 40, 57, 64]
        mats = []
        hosts = []
        for i, n in enumerate(ns):
            A = oracle.gen_reference_symmetric(n, 777 * n + bw + i)
            T = torch.zeros((n, n + 1), dtype=torch.float64, device="cuda")
            T[:, :n] = torch.from_numpy(np.triu(A)).cuda()
            mats.append(T)
            hosts.append(A)
        outs = factor_batch(handle, oracle, mats, torch.float64)
        checked = 0
        for A, g in zip(hosts, outs):
            o = check(oracle, A, g, bw)
            checked += int(oracle.min_margin(o.trace) > MARGIN)
        assert checked >= 3
        assert any((g.piv == 2).any() for g in outs)
    finally:
        handle.set_batched_widths(torch.float64, saved_w)
        handle.set_batched_team_sizes(torch.float64, saved_nt)


RESIDENT_NS = [1, 2, 3, 7, 16, 31, 32, 33, 47, 64, 65, 100, 127, 128, 129, 150, 176, 177, 200,
               230, 255, 256]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("bw,ts", [(4, 16), (16, 32), (32, 16), (24, 32)])
def test_resident_kernel_matches_oracle(oracle, handle, precision, bw, ts):
    """Batches whose largest matrix has n <= 256 run resident: the packed
    upper triangle in LDS (binary64 matrices above the LDS budget keep their
    first rows in global memory and defer those rows' interchanges), MFMA
    trailing updates with 16 x 16 or 32 x 32 tiles per wavefront."""
    import torch
    dtype = torch.float64 if precision == "fp64" else torch.float32
    saved_w, saved_t = handle.batched_widths(dtype), handle.batched_tiles(dtype)
    saved_nt = handle.batched_team_sizes(dtype)
    handle.set_batched_widths(dtype, [bw] * 8)
    handle.set_batched_tiles(dtype, [ts] * 8)
    handle.set_batched_team_sizes(dtype, [0, 0, 0, 0, 256, 256, 256, 256])
    try:
        mats = handle.gen_synth_batched(RESIDENT_NS, seed=7000 + bw + ts, scale=1.0,
                                        zero_diag_stride=7, dtype=dtype, pad=3)
        hosts = [host_matrix(m) for m in mats]
        outs = factor_batch(handle, oracle, mats, dtype)
        for A, g in zip(hosts, outs):
            check(oracle, A, g, bw, fp32=(precision == "fp32"))
        assert any((g.piv == 2).any() for g in outs)
    finally:
        handle.set_batched_widths(dtype, saved_w)
        handle.set_batched_tiles(dtype, saved_t)
        handle.set_batched_team_sizes(dtype, saved_nt)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("nt,ts", [(64, 16), (64, 32), (256, 16), (256, 32)])
def test_global_kernel_mfma_update(oracle, handle, precision, nt, ts):
    """The global-memory kernel with the MFMA trailing update (16 x 16 or
    32 x 32 tiles per wavefront) instead of SIMT tiles, one- and
    four-wavefront teams, orders up to 400 (W in LDS and in global slices)."""
    import torch
    dtype = torch.float64 if precision == "fp64" else torch.float32
    saved_t, saved_nt = handle.batched_tiles(dtype), handle.batched_team_sizes(dtype)
    handle.set_batched_tiles(dtype, [ts] * 8)
    handle.set_batched_team_sizes(dtype, [nt] * 8)
    try:
        ns = RESIDENT_NS + [300, 400]
        mats = handle.gen_synth_batched(ns, seed=9100 + nt + ts, scale=1.0, zero_diag_stride=7,
                                        dtype=dtype, pad=2)
        hosts = [host_matrix(m) for m in mats]
        widths = handle.batched_widths(dtype)
        outs = factor_batch(handle, oracle, mats, dtype)
        for A, g in zip(hosts, outs):
            check(oracle, A, g, width_for(widths, A.shape[0]), fp32=(precision == "fp32"))
    finally:
        handle.set_batched_tiles(dtype, saved_t)
        handle.set_batched_team_sizes(dtype, saved_nt)


def test_batched_global_workspace_path(oracle, handle):
    """Matrices whose panel rows exceed the LDS budget use global slices."""
    import torch
    ns = [700, 1100, 333]
    mats = handle.gen_synth_batched(ns, seed=99, scale=1e-6, zero_diag_stride=7,
                                    dtype=torch.float64)
    hosts = [host_matrix(m) for m in mats]
    widths = handle.batched_widths(torch.float64)
    outs = factor_batch(handle, oracle, mats, torch.float64)
    for A, g in zip(hosts, outs):
        check(oracle, A, g, width_for(widths, A.shape[0]))


@pytest.mark.parametrize("nt", [64, 256])
def test_batched_fp32(oracle, handle, nt):
    import torch
    saved_nt = handle.batched_team_sizes(torch.float32)
    handle.set_batched_team_sizes(torch.float32, [nt] * 8)
    ns = [20, 64, 100, 180, 256, 400]
    mats = handle.gen_synth_batched(ns, seed=31337, scale=1.0, zero_diag_stride=7,
                                    dtype=torch.float32)
    hosts = [host_matrix(m) for m in mats]  # the binary32 values, widened
    widths = handle.batched_widths(torch.float32)
    try:
        outs = factor_batch(handle, oracle, mats, torch.float32)
        for A, g in zip(hosts, outs):
            check(oracle, A, g, width_for(widths, A.shape[0]), fp32=True)
    finally:
        handle.set_batched_team_sizes(torch.float32, saved_nt)


def test_batched_equals_single_matrix_path(oracle, handle):
    """Same decisions as bbk_dsytrf on the same input (any panel width gives
    the same Bunch-Kaufman decisions on well-separated inputs)."""
    import torch
    ns = [150, 256, 333]
    mats = handle.gen_synth_batched(ns, seed=4242, zero_diag_stride=7, dtype=torch.float64)
    singles = [m.clone() for m in mats]
    outs = factor_batch(handle, oracle, mats, torch.float64)
    for T, g in zip(singles, outs):
        A = host_matrix(T)
        perm, piv = handle.sytrf(T)
        handle.synchronize()
        o = oracle.dsytrf(A, 32, trace=True)
        if oracle.min_margin(o.trace) > MARGIN:
            assert np.array_equal(piv.cpu().numpy(), g.piv)
            assert np.array_equal(perm.cpu().numpy(), g.perm)


def test_batched_empty_and_repeat(oracle, handle):
    import torch
    from blockedbunchkaufmanldlt_amd import lib
    assert lib().bbk_dsytrf_batched(handle.h, 0, None, None, None, None, None) == 0
    ns = [50, 60]
    base = handle.gen_synth_batched(ns, seed=1, zero_diag_stride=7, dtype=torch.float64)
    first = factor_batch(handle, oracle, [b.clone() for b in base], torch.float64)
    second = factor_batch(handle, oracle, [b.clone() for b in base], torch.float64)
    for a, b in zip(first, second):
        assert np.array_equal(a.compact, b.compact)
        assert np.array_equal(a.piv, b.piv)


def test_batched_graph_replay(oracle, handle):
    """A captured graph replays the factorization; every replay (inputs
    restored eagerly in between) equals the eager result."""
    import torch
    ns = [40, 90, 130, 200, 256, 64]
    src = handle.gen_synth_batched(ns, seed=2024, zero_diag_stride=7, dtype=torch.float64)
    eager = factor_batch(handle, oracle, [s.clone() for s in src], torch.float64)
    work = [s.clone() for s in src]
    perms = [torch.empty(n, dtype=torch.int32, device="cuda") for n in ns]
    pivs = [torch.empty(n, dtype=torch.int32, device="cuda") for n in ns]
    g = handle.graph_sytrf_batched(work, perms, pivs)
    try:
        for rep in range(3):
            for w, s in zip(work, src):
                w.copy_(s)
            torch.cuda.synchronize()
            g.launch()
            handle.synchronize()
            for w, p, v, e in zip(work, perms, pivs, eager):
                n = w.shape[0]
                assert np.array_equal(np.triu(w[:, :n].cpu().numpy()), e.compact), rep
                assert np.array_equal(v.cpu().numpy(), e.piv)
                assert np.array_equal(p.cpu().numpy(), e.perm)
    finally:
        g.close()


def test_batched_autotuner(oracle, handle):
    import torch
    saved = handle.batched_widths(torch.float64)
    saved_nt = handle.batched_team_sizes(torch.float64)
    try:
        widths = handle.batched_tune(torch.float64, 128)
        assert all(w in (8, 12, 16, 24, 32, 48, 64) for w in widths[:3])
        assert all(t in (0, 16, 32) for t in handle.batched_tiles(torch.float64))
        assert widths == handle.batched_widths(torch.float64)
        assert all(t in (0, 1, 64, 256) for t in handle.batched_team_sizes(torch.float64))
        ns = [30, 60, 120]
        mats = handle.gen_synth_batched(ns, seed=8, zero_diag_stride=7, dtype=torch.float64)
        hosts = [host_matrix(m) for m in mats]
        outs = factor_batch(handle, oracle, mats, torch.float64)
        for A, g in zip(hosts, outs):
            check(oracle, A, g, width_for(widths, A.shape[0]))
    finally:
        handle.set_batched_widths(torch.float64, saved)
        handle.set_batched_team_sizes(torch.float64, saved_nt)


@pytest.mark.parametrize("bw", [3, 32])
def test_batched_lockstep_matches_oracle(oracle, handle, bw):
    """Batches with a matrix above 512 run in lockstep (one panel launch +
    one batched MFMA update launch per panel step); tiny matrices in the same
    batch finish early; narrow panels defer many 2x2 pivots."""
    import torch
    saved = handle.batched_widths(torch.float64)
    handle.set_batched_widths(torch.float64, [bw] * 8)
    try:
        ns = [600, 1, 2, 17, 64, 300, 533]
        mats = handle.gen_synth_batched(ns, seed=777 + bw, scale=1.0, zero_diag_stride=7,
                                        dtype=torch.float64, pad=5)
        hosts = [host_matrix(m) for m in mats]
        outs = factor_batch(handle, oracle, mats, torch.float64)
        for A, g in zip(hosts, outs):
            check(oracle, A, g, bw)
        assert any((g.piv == 2).any() for g in outs)
    finally:
        handle.set_batched_widths(torch.float64, saved)


def test_batched_lockstep_fp32_and_graph(oracle, handle):
    import torch
    ns = [700, 520, 90, 256]
    src = handle.gen_synth_batched(ns, seed=99, zero_diag_stride=7, dtype=torch.float32)
    eager = factor_batch(handle, oracle, [s.clone() for s in src], torch.float32)
    widths = handle.batched_widths(torch.float32)
    for s_, e in zip(src, eager):
        check(oracle, host_matrix(s_), e, width_for(widths, s_.shape[0]), fp32=True)
    work = [s.clone() for s in src]
    perms = [torch.empty(n, dtype=torch.int32, device="cuda") for n in ns]
    pivs = [torch.empty(n, dtype=torch.int32, device="cuda") for n in ns]
    g = handle.graph_sytrf_batched(work, perms, pivs)
    try:
        for rep in range(2):
            for w, s_ in zip(work, src):
                w.copy_(s_)
            torch.cuda.synchronize()
            g.launch()
            handle.synchronize()
            for w, p, v, e in zip(work, perms, pivs, eager):
                n = w.shape[0]
                assert np.array_equal(np.triu(w[:, :n].double().cpu().numpy()), e.compact), rep
                assert np.array_equal(v.cpu().numpy(), e.piv)
    finally:
        g.close()


@pytest.mark.parametrize("ns", [[17, 64, 130, 200, 256, 255, 1], [600, 1031, 17, 300, 513]])
def test_batched_factors_and_residuals_match_single_matrix_path(oracle, handle, ns):
    """Per matrix, the batched factorization (the symmetric-workspace kernel
    of the default tables for the first batch, lockstep launches for the
    second) against the single-matrix path bbk_dsytrf on the same input:
    identical pivot types and permutation whenever every decision is well
    separated (the two differ only in panel geometry and summation
    grouping), compact factors to 1e-9 relative, and device residuals within
    10x of each other."""
    import torch
    import blockedbunchkaufmanldlt_amd as B
    mats = handle.gen_synth_batched(ns, seed=5151 + len(ns), scale=1.0, zero_diag_stride=7,
                                    dtype=torch.float64, pad=3)
    srcs = [m.clone() for m in mats]
    perms = [torch.empty(n, dtype=torch.int32, device="cuda") for n in ns]
    pivs = [torch.empty(n, dtype=torch.int32, device="cuda") for n in ns]
    handle.sytrf_batched(mats, perms, pivs)
    handle.synchronize()
    separated = 0
    for m, src, p, v in zip(mats, srcs, perms, pivs):
        n = m.shape[0]
        T = torch.zeros((n, B.padded_ld(n)), dtype=torch.float64, device="cuda")
        T[:, :n] = src[:, :n]
        ps, vs = handle.sytrf(T)
        handle.synchronize()
        r_b = handle.residual(src, m, p, v)
        r_s = handle.residual(src, T, ps, vs)
        assert r_b < 1e-10 and r_s < 1e-10, (n, r_b, r_s)
        assert r_b <= 10 * max(r_s, 1e-15) and r_s <= 10 * max(r_b, 1e-15), (n, r_b, r_s)
        o = oracle.dsytrf(host_matrix(src), 64, trace=True)
        if oracle.min_margin(o.trace) > MARGIN:
            separated += 1
            assert torch.equal(p, ps), n
            assert torch.equal(v, vs), n
            Fb = np.triu(m[:, :n].cpu().numpy())
            Fs = np.triu(T[:, :n].cpu().numpy())
            scale = max(1.0, np.abs(Fs).max())
            assert np.abs(Fb - Fs).max() / scale < 1e-9, n
    assert separated >= len(ns) // 2

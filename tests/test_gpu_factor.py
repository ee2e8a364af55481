"""GPU parity tests for the plain factorization (bbk_dsytrf / bbk_ssytrf).

The oracle (oracle/bk_oracle.c, bitwise equal to the reference) is the
checker.  Bars, from the reference's own tests and the FP64 target:
* pivot types and permutation identical to the oracle whenever every
  decision is well separated (relative margin > 1e-8);
* reconstruction residual ||P L D L^T P^T - A||_F / max(||A||_F, 1)
  < 1e-10 (tests/block_ldlt_decomposition_test.cc:31) and within 10x of the
  reference's residual on the same input;
* compact factors equal to the oracle's to 1e-9 relative when the pivot
  sequences agree;
* sampled residual < 1e-8 at n = 5000 / 10000
  (tests/block_ldlt_large_stress_test.cc:26).
"""
import os

import numpy as np
import pytest

from gpu_util import gpu_factor, to_device

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_factors.npz")
MARGIN = 1e-8


def check_against_oracle(oracle, A, g, bs, res_floor=1e-15):
    o = oracle.dsytrf(A, bs, trace=True)
    res_g = oracle.relative_residual(A, g)
    res_o = oracle.relative_residual(A, o)
    assert res_g < 1e-10
    assert res_g <= 10.0 * max(res_o, res_floor)
    if oracle.min_margin(o.trace) > MARGIN:
        assert np.array_equal(g.piv, o.piv)
        assert np.array_equal(g.perm, o.perm)
        scale = max(1.0, np.abs(o.compact).max())
        assert np.abs(g.compact - o.compact).max() / scale < 1e-9
    return res_g, res_o


@pytest.mark.parametrize("n,nb", [(1, 1), (2, 1), (3, 2), (4, 2), (5, 3), (8, 1), (16, 4),
                                  (17, 8), (33, 5), (64, 16), (100, 32), (129, 64),
                                  (257, 64), (300, 0), (513, 48), (700, 64),
                                  (200, 100), (513, 128), (1100, 0), (1500, 120)])
def test_random_matches_oracle(oracle, handle, n, nb):
    for seed in range(3):
        A = oracle.gen_reference_symmetric(n, 10000 * n + 100 * nb + seed)
        g = gpu_factor(handle, oracle, A, nb)
        check_against_oracle(oracle, A, g, max(nb, 1))


def test_reference_kat(oracle, handle):
    """TestReferenceCompactStorage on the GPU path."""
    A = np.array([[6.0, 12.0, 3.0, -6.0], [12.0, -8.0, -13.0, 4.0],
                  [3.0, -13.0, -1.0, 1.0], [-6.0, 4.0, 1.0, 6.0]])
    expected = np.array([[6.0, 12.0, -0.6875, 0.0], [0.0, -8.0, 0.59375, -0.5],
                         [0.0, 0.0, 8.78125, -0.6263345195729537],
                         [0.0, 0.0, 0.0, 4.555160142348754]])
    for nb in (1, 2, 3, 4):
        g = gpu_factor(handle, oracle, A, nb)
        assert list(g.piv) == [2, -2, 1, 1]
        assert np.linalg.norm(g.compact - expected) / np.linalg.norm(expected) < 1e-10


def test_golden_reference_outputs(oracle, handle):
    """Against the reference's own outputs committed as golden vectors."""
    g = np.load(GOLDEN)
    for idx, (n, bs, seed, scale) in enumerate(g["meta"]):
        n, bs = int(n), int(bs)
        iu = np.triu_indices(n)
        A = np.zeros((n, n))
        A[iu] = g["A_%d" % idx]
        A = A + np.triu(A, 1).T
        f = gpu_factor(handle, oracle, A, bs)
        res_ref = g["res_%d" % idx][0]
        res = oracle.relative_residual(A, f)
        assert res < 1e-10 and res <= 10 * max(res_ref, 1e-15), idx
        o = oracle.dsytrf(A, bs, trace=True)
        if oracle.min_margin(o.trace) > MARGIN:
            assert np.array_equal(f.piv, g["piv_%d" % idx]), idx
            assert np.array_equal(f.perm, g["perm_%d" % idx]), idx
            U = g["U_%d" % idx]
            assert np.abs(f.compact[iu] - U).max() <= 1e-9 * max(1.0, np.abs(U).max()), idx


@pytest.mark.parametrize("case", ["pairs", "singular", "zero", "scaled_small",
                                  "scaled_large", "zero_diag7", "dominant"])
def test_edge_cases(oracle, handle, case):
    n = 96
    rng = np.random.default_rng(3)
    if case == "pairs":  # forced 2x2 pivots (tests/...:485-491)
        A = np.zeros((n, n))
        for k in range(0, n, 2):
            A[k, k + 1] = A[k + 1, k] = 1.0 + k
    elif case == "singular":  # diag(2, -3, 0, 5) pattern
        A = np.diag(np.tile([2.0, -3.0, 0.0, 5.0], n // 4))
    elif case == "zero":
        A = np.zeros((n, n))
    elif case == "scaled_small":
        A = oracle.gen_synth_symmetric(n, 11, 1e-12)
    elif case == "scaled_large":
        A = oracle.gen_synth_symmetric(n, 12, 1e12)
    elif case == "zero_diag7":
        A = oracle.gen_synth_symmetric(n, 13, 1.0, 7)
    else:
        A = rng.uniform(-1, 1, (n, n)); A = A + A.T
        A[np.diag_indices(n)] = 3 * n
    for nb in (1, 5, 32, 128):
        g = gpu_factor(handle, oracle, A, nb)
        assert np.all(np.isfinite(g.compact))
        check_against_oracle(oracle, A, g, nb, res_floor=1e-15)
    if case == "pairs":
        assert list(g.piv) == [2, -2] * (n // 2)
    if case == "zero":
        assert np.all(g.compact == 0) and np.all(g.piv == 1)


def test_deterministic(oracle, handle):
    A = oracle.gen_synth_symmetric(777, 21)
    a = gpu_factor(handle, oracle, A, 64)
    b = gpu_factor(handle, oracle, A, 64)
    assert np.array_equal(a.compact, b.compact) and np.array_equal(a.perm, b.perm)


def test_gen_synth_bitwise(oracle, handle):
    n = 300
    T = handle.gen_synth(n, 12345, 3.5, 7, lda=320)
    handle.synchronize()
    assert np.array_equal(T[:, :n].cpu().numpy(), oracle.gen_synth_symmetric(n, 12345, 3.5, 7))


@pytest.mark.parametrize("n,scale", [(5000, 1.0), (5000, 1e-6), (10000, 1e6)])
def test_large_sampled_residual(oracle, handle, n, scale):
    """Config: plain FP64 drop-in, block size 128, upper entries
    uniform[-1,1] x value_scale, zero diagonal on i % 7 == 0; the sampled
    residual of tests/block_ldlt_large_stress_test.cc:220-244 (4096 samples)."""
    import torch
    import blockedbunchkaufmanldlt_amd as B
    seed = 930000 + n
    lda = B.padded_ld(n)
    T = handle.gen_synth(n, seed, scale, 7, lda=lda)
    perm, piv = handle.sytrf(T, 128)
    handle.synchronize()
    F = np.triu(T[:, :n].cpu().numpy())
    del T
    torch.cuda.empty_cache()
    A = oracle.gen_synth_symmetric(n, seed, scale, 7)
    f = oracle.Factorization(n, 128, F, perm.cpu().numpy().astype(np.int32),
                             piv.cpu().numpy().astype(np.int32))
    assert np.all(np.isfinite(F))
    res = oracle.sampled_residual(A, f, seed)
    assert res < 1e-8
    assert sorted(f.perm) == list(range(n))


def test_large_pivots_match_oracle(oracle, handle):
    """Full-size decision parity: n = 5000 against the OpenMP oracle."""
    n = 5000
    seed = 931300
    A = oracle.gen_synth_symmetric(n, seed, 1.0, 7)
    g = gpu_factor(handle, oracle, A, 128)
    o = oracle.dsytrf(A, 128, trace=True)
    assert np.array_equal(g.piv, o.piv)
    assert np.array_equal(g.perm, o.perm)
    rows, cols = oracle.residual_samples(n, seed)
    rg = oracle.sampled_residual(A, g, seed)
    ro = oracle.sampled_residual(A, o, seed)
    assert rg < 1e-8 and rg <= 10 * max(ro, 1e-15)


@pytest.mark.parametrize("n", [64, 300, 1000])
def test_fp32_factorization(oracle, handle, n):
    """FP32 form: same decisions as the FP64 oracle on well-separated
    matrices; residual at binary32 level."""
    A = oracle.gen_synth_symmetric(n, 500 + n)
    g = gpu_factor(handle, oracle, A, 0, dtype="float32")
    o = oracle.dsytrf(A, 64, trace=True)
    res = oracle.relative_residual(A, g)
    assert res < 1e-4
    if oracle.min_margin(o.trace) > 1e-3:
        assert np.array_equal(g.piv, o.piv)
        assert np.array_equal(g.perm, o.perm)


def test_kernels_shapes(oracle, handle):
    """TestKernelShapes (tests/block_ldlt_decomposition_test.cc:367-385) and
    random strided cases of the kernels.h counterparts against the oracle."""
    import torch
    dev = "cuda"
    src = torch.tensor([2.0, 4.0, 6.0], device=dev, dtype=torch.float64)
    dst = torch.zeros(3, device=dev, dtype=torch.float64)
    handle.kernel_scale_1x1(3, 2.0, src, 1, dst, 1)
    s1 = torch.tensor([1.0, 2.0], device=dev, dtype=torch.float64)
    s2 = torch.tensor([3.0, 4.0], device=dev, dtype=torch.float64)
    d1 = torch.zeros(2, device=dev, dtype=torch.float64)
    d2 = torch.zeros(2, device=dev, dtype=torch.float64)
    handle.kernel_scale_2x2(2, 2.0, 1.0, 3.0, s1, 1, s2, 1, d1, 1, d2, 1)
    handle.synchronize()
    assert dst[0].item() == 1.0 and dst[2].item() == 3.0
    assert abs(d1[0].item()) < 1e-14 and abs(d2[0].item() - 1.0) < 1e-14
    rng = np.random.default_rng(0)
    L = oracle.lib()
    # update_trailing: rows x cols target, inner rows of operands, strides
    for (rows, cols, inner) in [(1, 1, 1), (17, 33, 5), (64, 64, 64), (130, 70, 129),
                                (200, 257, 128), (65, 300, 9)]:
        ls, ws, ts = rows + 3, cols + 1, cols + 2
        left = rng.standard_normal(inner * ls)
        work = rng.standard_normal(inner * ws)
        tgt = rng.standard_normal(rows * ts)
        want = tgt.copy()
        L.bko_kernel_update_trailing(rows, cols, inner, oracle._d(left), ls, oracle._d(work),
                                     ws, oracle._d(want), ts)
        tl, tw, tt = (torch.from_numpy(x).to(dev) for x in (left, work, tgt))
        handle.kernel_update_trailing(rows, cols, inner, tl, ls, tw, ws, tt, ts)
        handle.synchronize()
        assert np.allclose(tt.cpu().numpy(), want, rtol=1e-12, atol=1e-12)
    # scale kernels with strides, including exact-zero pivots
    for pivot in (0.0, -3.25):
        s = rng.standard_normal(40)
        want = np.zeros(40)
        L.bko_kernel_scale_1x1(20, pivot, oracle._d(s), 2, oracle._d(want), 2)
        ts_, td = torch.from_numpy(s).to(dev), torch.zeros(40, dtype=torch.float64, device=dev)
        handle.kernel_scale_1x1(20, pivot, ts_, 2, td, 2)
        handle.synchronize()
        assert np.array_equal(td.cpu().numpy(), want)
    for (a, b, c) in ((2.0, 1.0, 3.0), (1.0, 1.0, 1.0), (-1.5, 4.0, 0.25)):
        x1, x2 = rng.standard_normal(30), rng.standard_normal(30)
        w1, w2 = np.zeros(30), np.zeros(30)
        L.bko_kernel_scale_2x2(10, a, b, c, oracle._d(x1), 3, oracle._d(x2), 3,
                               oracle._d(w1), 3, oracle._d(w2), 3)
        t1, t2 = torch.from_numpy(x1).to(dev), torch.from_numpy(x2).to(dev)
        o1, o2 = torch.zeros(30, dtype=torch.float64, device=dev), \
            torch.zeros(30, dtype=torch.float64, device=dev)
        handle.kernel_scale_2x2(10, a, b, c, t1, 3, t2, 3, o1, 3, o2, 3)
        handle.synchronize()
        assert np.array_equal(o1.cpu().numpy(), w1) and np.array_equal(o2.cpu().numpy(), w2)


def test_kernels_host_forms(oracle, handle):
    """The host-pointer forms that the drop-in's naive_block_ldlt::kernels::*
    call (operands copied to the device and back) against the oracle's
    restatement of src/kernels.cc, with strides and an exact-zero pivot."""
    import ctypes
    import blockedbunchkaufmanldlt_amd as B
    lib = B.lib()
    vp, i32, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.bbk_kernel_scale_1x1_host.argtypes = [vp, i32, f64, vp, i32, vp, i32]
    lib.bbk_kernel_scale_2x2_host.argtypes = [vp, i32, f64, f64, f64, vp, i32, vp, i32, vp, i32,
                                              vp, i32]
    lib.bbk_kernel_update_trailing_host.argtypes = [vp, i32, i32, i32, vp, i32, vp, i32, vp, i32]
    L = oracle.lib()
    d = oracle._d
    rng = np.random.default_rng(5)
    for (rows, cols, inner) in [(1, 1, 1), (19, 7, 3), (70, 130, 128)]:
        ls, ws, ts = rows + 2, cols + 3, cols + 1
        left = rng.standard_normal(inner * ls)
        work = rng.standard_normal(inner * ws)
        tgt = rng.standard_normal(rows * ts)
        want, got = tgt.copy(), tgt.copy()
        L.bko_kernel_update_trailing(rows, cols, inner, d(left), ls, d(work), ws, d(want), ts)
        assert lib.bbk_kernel_update_trailing_host(handle.h, rows, cols, inner, d(left), ls,
                                                   d(work), ws, d(got), ts) == 0
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
    for pivot in (0.0, 1.75):
        s = rng.standard_normal(33)
        want, got = np.zeros(33), np.zeros(33)
        L.bko_kernel_scale_1x1(11, pivot, d(s), 3, d(want), 3)
        assert lib.bbk_kernel_scale_1x1_host(handle.h, 11, pivot, d(s), 3, d(got), 3) == 0
        assert np.array_equal(got, want)
    x1, x2 = rng.standard_normal(24), rng.standard_normal(24)
    w1, w2, g1, g2 = np.zeros(24), np.zeros(24), np.zeros(24), np.zeros(24)
    L.bko_kernel_scale_2x2(12, -1.5, 4.0, 0.25, d(x1), 2, d(x2), 2, d(w1), 2, d(w2), 2)
    assert lib.bbk_kernel_scale_2x2_host(handle.h, 12, -1.5, 4.0, 0.25, d(x1), 2, d(x2), 2,
                                         d(g1), 2, d(g2), 2) == 0
    assert np.array_equal(g1, w1) and np.array_equal(g2, w2)

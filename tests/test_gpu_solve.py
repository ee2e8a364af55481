"""GPU tests of the solve, backward-error and mixed-precision refinement
entry points (no reference counterpart; pinned by the oracle's solve,
numpy's dense solve and the Rigal-Gaches backward-error formula)."""
import numpy as np
import pytest

from gpu_util import to_device

pytestmark = pytest.mark.gpu


def eig_matrix(n, kappa, seed):
    """A = Q diag(lambda) Q^T, |lambda| log-spaced over [1/kappa, 1], random signs."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = 10.0 ** (-np.log10(kappa) * np.arange(n) / max(n - 1, 1))
    lam *= np.where(rng.random(n) < 0.5, -1.0, 1.0)
    A = (Q * lam) @ Q.T
    return 0.5 * (A + A.T)


def berr_np(A, B, X):
    R = B - A @ X
    return (np.abs(R).max(0) / (np.abs(A).sum(1).max() * np.abs(X).max(0) + np.abs(B).max(0)))


@pytest.mark.parametrize("n,nrhs", [(1, 1), (7, 2), (64, 1), (65, 3), (200, 5), (1000, 4)])
def test_dsytrs(oracle, handle, n, nrhs):
    import torch
    A = oracle.gen_synth_symmetric(n, 40 + n)
    T = to_device(A)
    perm, piv = handle.sytrf(T, 0)
    rng = np.random.default_rng(n)
    Bh = rng.standard_normal((n, nrhs))
    Bd = torch.from_numpy(Bh.copy()).cuda()
    handle.sytrs(T, perm, piv, Bd)
    handle.synchronize()
    X = Bd.cpu().numpy()
    # against the oracle's solve on the GPU's own factors
    F = np.triu(T[:, :n].cpu().numpy())
    f = oracle.Factorization(n, 0, F, perm.cpu().numpy().astype(np.int32),
                             piv.cpu().numpy().astype(np.int32))
    Xo = oracle.dsytrs(f, Bh)
    assert np.allclose(X, Xo, rtol=1e-10, atol=1e-11)
    assert berr_np(A, Bh, X).max() < 1e-14


@pytest.mark.parametrize("n", [100, 777])
def test_ssytrs(oracle, handle, n):
    import torch
    A = oracle.gen_synth_symmetric(n, 90 + n)
    T = to_device(A, "float32")
    perm, piv = handle.sytrf(T, 0)
    Bh = np.random.default_rng(2).standard_normal((n, 2))
    Bd = torch.from_numpy(Bh.astype(np.float32)).cuda()
    handle.sytrs(T, perm, piv, Bd)
    handle.synchronize()
    X = Bd.double().cpu().numpy()
    assert berr_np(A, Bh, X).max() < 1e-5


def test_dsyberr(oracle, handle):
    import torch
    n = 300
    A = oracle.gen_synth_symmetric(n, 5)
    rng = np.random.default_rng(9)
    B = rng.standard_normal((n, 6))
    X = np.linalg.solve(A, B) + 1e-9 * rng.standard_normal((n, 6))
    eta = handle.syberr(to_device(A), torch.from_numpy(B).cuda(), torch.from_numpy(X).cuda())
    handle.synchronize()
    assert np.allclose(eta.cpu().numpy(), berr_np(A, B, X), rtol=1e-10)


@pytest.mark.parametrize("n,kappa", [(256, 1e4), (1500, 1e6)])
def test_dsysv_fp64_and_mixed(oracle, handle, n, kappa):
    import torch
    A = eig_matrix(n, kappa, 7 + n)
    b = np.random.default_rng(11).standard_normal((n, 1))
    Ad = to_device(A)
    Ad[:, :n] = torch.from_numpy(A).cuda()  # full matrix; only the upper is read
    bd = torch.from_numpy(b).cuda()
    X = torch.zeros_like(bd)
    be64 = handle.sysv(Ad, bd, X)
    x64 = X.cpu().numpy()
    assert be64 < 1e-14 and berr_np(A, b, x64).max() < 1e-14
    X.zero_()
    info = handle.sysv_mixed(Ad, bd, X, tol=1e-12, max_iter=30)
    xm = X.cpu().numpy()
    assert info["converged"] and info["berr"] <= 1e-12
    assert berr_np(A, b, xm).max() <= 1.5e-12
    # forward error bounded by condition number times backward error
    xr = np.linalg.solve(A, b)
    assert np.abs(xm - xr).max() / np.abs(xr).max() < kappa * 1e-11


def test_mixed_reports_non_convergence(oracle, handle):
    """Condition 1e12 is beyond what FP32 factors can refine: the driver must
    say so rather than return a silently wrong answer."""
    import torch
    n = 200
    A = eig_matrix(n, 1e12, 3)
    b = np.random.default_rng(4).standard_normal((n, 1))
    Ad = to_device(A)
    Ad[:, :n] = torch.from_numpy(A).cuda()
    X = torch.zeros((n, 1), dtype=torch.float64, device="cuda")
    info = handle.sysv_mixed(Ad, torch.from_numpy(b).cuda(), X, tol=1e-15, max_iter=3)
    assert not info["converged"] and info["iters"] == 3
    # X holds the best iterate, whose backward error is the one reported
    Xh = X.cpu().numpy()
    assert np.isfinite(Xh).all()
    be = berr_np(A, b, Xh).max()
    assert abs(be - info["berr"]) <= 1e-6 * be, (be, info["berr"])


def test_mixed_returns_best_iterate_when_refinement_fails(oracle, handle):
    """Condition 1e13: binary32 corrections cannot reach tol and may make x
    worse.  The driver must return the best iterate (never worse than x0)
    with its backward error, not the last one."""
    import torch
    n = 300
    A = eig_matrix(n, 1e13, 17)
    b = np.random.default_rng(18).standard_normal((n, 1))
    Ad = to_device(A)
    Ad[:, :n] = torch.from_numpy(A).cuda()
    bd = torch.from_numpy(b).cuda()
    X0 = torch.zeros((n, 1), dtype=torch.float64, device="cuda")
    info0 = handle.sysv_mixed(Ad, bd, X0, tol=1e-15, max_iter=0)
    be0 = berr_np(A, b, X0.cpu().numpy()).max()
    X = torch.zeros((n, 1), dtype=torch.float64, device="cuda")
    info = handle.sysv_mixed(Ad, bd, X, tol=1e-15, max_iter=12)
    assert not info["converged"]
    Xh = X.cpu().numpy()
    assert np.isfinite(Xh).all()
    be = berr_np(A, b, Xh).max()
    assert abs(be - info["berr"]) <= 1e-6 * be, (be, info["berr"])
    assert be <= be0 * (1 + 1e-9), (be, be0, info0["berr"])


def test_multi_rhs_mixed(oracle, handle):
    """Batched-config shape: 4 right-hand sides refined together."""
    import torch
    n = 512
    A = eig_matrix(n, 1e5, 21)
    B = np.random.default_rng(22).standard_normal((n, 4))
    Ad = to_device(A)
    Ad[:, :n] = torch.from_numpy(A).cuda()
    X = torch.zeros((n, 4), dtype=torch.float64, device="cuda")
    info = handle.sysv_mixed(Ad, torch.from_numpy(B).cuda(), X, tol=1e-12)
    assert info["converged"]
    assert berr_np(A, B, X.cpu().numpy()).max() <= 1.5e-12


@pytest.mark.parametrize("n", [1000, 1153])
def test_mixed_many_2x2_ragged(oracle, handle, n):
    """Refinement whose FP32 solves multiply by inverted 128 x 128 diagonal
    blocks: orders that are not multiples of 128 and a zero diagonal (2x2
    pivots everywhere, also across block edges).  A wrong block inverse or
    edge entry stalls the refinement, so the iteration count is bounded."""
    import torch
    rng = np.random.default_rng(n)
    A = rng.uniform(-1.0, 1.0, (n, n))
    A = np.triu(A) + np.triu(A, 1).T
    np.fill_diagonal(A, 0.0)
    b = rng.standard_normal((n, 2))
    Ad = to_device(A)
    Ad[:, :n] = torch.from_numpy(A).cuda()
    X = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    info = handle.sysv_mixed(Ad, torch.from_numpy(b).cuda(), X, tol=1e-12, max_iter=30)
    assert info["converged"] and info["iters"] <= 8, info
    assert berr_np(A, b, X.cpu().numpy()).max() <= 1.5e-12


def test_mixed_six_rhs_two_chunks(oracle, handle):
    """Six right-hand sides: the refinement's sweeps run in chunks of four
    and two, all under one device-side stopping test."""
    import torch
    n = 700
    A = eig_matrix(n, 1e5, 31)
    B6 = np.random.default_rng(32).standard_normal((n, 6))
    Ad = to_device(A)
    Ad[:, :n] = torch.from_numpy(A).cuda()
    X = torch.zeros((n, 6), dtype=torch.float64, device="cuda")
    info = handle.sysv_mixed(Ad, torch.from_numpy(B6).cuda(), X, tol=1e-12)
    assert info["converged"] and info["berr"] <= 1e-12
    be = berr_np(A, B6, X.cpu().numpy())
    assert be.max() <= 1.5e-12 and abs(be.max() - info["berr"]) <= 1e-13


def test_mixed_refines_like_lapack_ssytrf(oracle, handle):
    """The binary32 factor is as accurate as LAPACK's: classic refinement
    with it needs no more corrections than the same refinement driven by
    scipy's ssytrf/ssytrs factor of the same matrix (condition 1e6, where
    every correction gains only about one decimal digit, so an inaccurate
    factor shows up as extra corrections)."""
    import torch
    from scipy.linalg import lapack
    n = 3000
    A = eig_matrix(n, 1e6, 41)
    b = np.random.default_rng(42).standard_normal(n)
    lu, ipiv, info = lapack.ssytrf(A.astype(np.float32), lower=1)
    assert info == 0

    def minv(r):
        return lapack.ssytrs(lu, ipiv, r.astype(np.float32), lower=1)[0].astype(np.float64)

    x = minv(b)
    lapack_iters = 0
    while berr_np(A, b[:, None], x[:, None]).max() > 1e-12 and lapack_iters < 30:
        x = x + minv(b - A @ x)
        lapack_iters += 1
    assert 2 <= lapack_iters < 30
    Ad = to_device(A)
    Ad[:, :n] = torch.from_numpy(A).cuda()
    X = torch.zeros((n, 1), dtype=torch.float64, device="cuda")
    res = handle.sysv_mixed(Ad, torch.from_numpy(b[:, None].copy()).cuda(), X, tol=1e-12,
                            max_iter=30)
    assert res["converged"]
    assert res["iters"] <= lapack_iters, (res["iters"], lapack_iters)


@pytest.mark.parametrize("n,nrhs,lda", [(1, 1, 1), (5, 3, 5), (63, 2, 67), (65, 4, 65),
                                        (130, 1, 131)])
def test_dsyberr_small_and_ragged(oracle, handle, n, nrhs, lda):
    """Backward error on orders below, at and just past one 64 x 64 residual
    tile, with unpadded leading dimensions; right-hand sides 1 to 4 (every
    instantiation of the tile kernel)."""
    import torch
    A = oracle.gen_synth_symmetric(n, 300 + n)
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, nrhs))
    X = rng.standard_normal((n, nrhs))
    Ad = to_device(A, lda=lda)
    eta = handle.syberr(Ad, torch.from_numpy(B).cuda(), torch.from_numpy(X).cuda())
    handle.synchronize()
    assert np.allclose(eta.cpu().numpy(), berr_np(A, B, X), rtol=1e-12)


@pytest.mark.parametrize("n,lda", [(77, 77), (200, 203), (256, 258)])
def test_solves_with_odd_leading_dimension(oracle, handle, n, lda):
    """Leading dimensions that are not multiples of four: the solves copy
    the input's upper triangle element by element instead of as vectors."""
    import torch
    A = eig_matrix(n, 1e4, 50 + n)
    b = np.random.default_rng(51).standard_normal((n, 1))
    Ad = to_device(A, lda=lda)
    bd = torch.from_numpy(b).cuda()
    X = torch.zeros_like(bd)
    be64 = handle.sysv(Ad, bd, X)
    assert be64 < 1e-14 and berr_np(A, b, X.cpu().numpy()).max() < 1e-14
    X.zero_()
    info = handle.sysv_mixed(Ad, bd, X, tol=1e-12, max_iter=30)
    assert info["converged"] and berr_np(A, b, X.cpu().numpy()).max() <= 1.5e-12

"""GPU tests of the LAPACK-convention interop (bbk_dsytrf_to_lapack,
bbk_dsytrf_from_lapack), pinned by LAPACK itself (scipy's dsytrs/dsytrf,
library routines, not reference code): a factor exported from this library
solves through LAPACK dsytrs, and a LAPACK factor imported into this
library solves through bbk_dsytrs and reconstructs A on the device."""
import numpy as np
import pytest

from gpu_util import to_device

pytestmark = pytest.mark.gpu
lapack = pytest.importorskip("scipy.linalg.lapack")


def sym(n, seed, zero_diag=False):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1.0, 1.0, (n, n))
    A = np.triu(A) + np.triu(A, 1).T
    if zero_diag:
        np.fill_diagonal(A, 0.0)
    return A


@pytest.mark.parametrize("n,nb,zero_diag", [(1, 0, False), (7, 2, True), (300, 32, False),
                                            (513, 64, True), (1000, 128, False)])
def test_export_solves_with_lapack_dsytrs(handle, n, nb, zero_diag):
    import torch
    A = sym(n, 100 + n, zero_diag)
    T = to_device(A)
    perm, piv = handle.sytrf(T, nb)
    ipiv = handle.to_lapack(T, perm, piv)
    handle.synchronize()
    # row-major upper triangle = column-major lower triangle
    a_lower = np.asfortranarray(np.tril(T[:, :n].cpu().numpy().T))
    ip = ipiv.cpu().numpy().astype(np.int32)
    if zero_diag and n > 1:
        assert np.any(ip < 0)  # a zero diagonal forces 2x2 pivots
    b = np.random.default_rng(7).standard_normal((n, 3))
    x, info = lapack.dsytrs(a_lower, ip, b, lower=1)
    assert info == 0
    xr = np.linalg.solve(A, b)
    assert np.abs(x - xr).max() <= 1e-8 * max(1.0, np.abs(xr).max())
    # and back: the round trip restores this library's factor bit for bit
    T2 = to_device(A)
    perm2, piv2 = handle.sytrf(T2, nb)
    p3, v3 = handle.from_lapack(T, ipiv)
    handle.synchronize()
    assert torch.equal(T[:, :n].triu(), T2[:, :n].triu())
    assert torch.equal(p3, perm2) and torch.equal(v3, piv2)


@pytest.mark.parametrize("n,zero_diag", [(5, True), (200, False), (777, True)])
def test_import_lapack_factor(handle, n, zero_diag):
    import torch
    A = sym(n, 300 + n, zero_diag)
    lu, ipiv, info = lapack.dsytrf(A, lower=1)
    assert info == 0
    # LAPACK's lower triangle (column-major) = this library's upper (row-major)
    T = to_device(np.triu(np.asarray(lu).T))
    ip = torch.from_numpy(np.asarray(ipiv, dtype=np.int32)).cuda()
    perm, piv = handle.from_lapack(T, ip)
    Ad = to_device(A)
    r = handle.residual(Ad, T, perm, piv)
    assert r < 1e-13, r
    b = np.random.default_rng(9).standard_normal((n, 2))
    X = torch.from_numpy(b.copy()).cuda()
    handle.sytrs(T, perm, piv, X)
    handle.synchronize()
    xr = np.linalg.solve(A, b)
    assert np.abs(X.cpu().numpy() - xr).max() <= 1e-8 * max(1.0, np.abs(xr).max())


def test_invalid_ipiv_is_rejected(handle):
    import torch
    n = 4
    T = to_device(np.eye(n))
    bad = torch.tensor([2, -3, 4, 4], dtype=torch.int32, device="cuda")  # lone negative
    import blockedbunchkaufmanldlt_amd as B
    with pytest.raises(B.BbkError):
        handle.from_lapack(T, bad)

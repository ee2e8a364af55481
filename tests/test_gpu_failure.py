"""GPU tests of failure reporting: the bounded waits of the cooperative
kernels and the device info word.

The wait-limit tests force the error from the host (a negative time limit
makes every wait of the panel exchanges and of the triangular sweeps fail
at once): the kernels still run to their end, the handle reports
BBK_ERR_HIP with the wait named, and the next call with the default limit
is correct again.  No kernel faults or hangs in these tests.

The info word follows LAPACK dsytrf's info: the first column whose 1x1
pivot is singular (|d| <= DBL_EPSILON, where the reference's
SafeDenominator zeroes the L column, src/block_ldlt_decomposition.cc:21-26),
plus a flag for NaN or infinite values in D (where a non-finite input
ends up).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _factor(handle, A, nb=0):
    import torch
    from blockedbunchkaufmanldlt_amd import padded_ld
    n = A.shape[0]
    T = torch.zeros((n, padded_ld(n)), dtype=torch.float64, device="cuda")
    T[:, :n] = torch.from_numpy(np.triu(A)).cuda()
    perm, piv = handle.sytrf(T, nb)
    return T, perm, piv


def test_panel_wait_limit_reports_error_then_recovers(oracle, handle, bbk):
    n = 3000
    A = oracle.gen_synth_symmetric(n, 4401, 1.0, 7)
    handle.set_wait_limit(-1)
    try:
        _factor(handle, A)
        with pytest.raises(bbk.BbkError) as ei:
            handle.synchronize()
        assert ei.value.code == bbk.BBK_ERR_HIP
        assert "panel exchange" in str(ei.value)
    finally:
        handle.set_wait_limit(0)
    # the error word was cleared; the next factorization is the usual one
    T, perm, piv = _factor(handle, A)
    handle.synchronize()
    F = np.triu(T[:, :n].cpu().numpy())
    g = oracle.Factorization(n, 0, F, perm.cpu().numpy().astype(np.int32),
                             piv.cpu().numpy().astype(np.int32))
    assert oracle.relative_residual(A, g) < 1e-10
    assert handle.info() == (0, 0)


def test_sweep_wait_limit_reports_error_then_recovers(oracle, handle, bbk):
    import torch
    n = 1000
    A = oracle.gen_synth_symmetric(n, 4402, 1.0, 7)
    T, perm, piv = _factor(handle, A)
    handle.synchronize()
    b = np.random.default_rng(1).standard_normal((n, 2))
    B = torch.from_numpy(b.copy()).cuda()
    handle.set_wait_limit(-1)
    try:
        handle.sytrs(T, perm, piv, B)
        with pytest.raises(bbk.BbkError) as ei:
            handle.synchronize()
        assert "triangular sweep" in str(ei.value)
    finally:
        handle.set_wait_limit(0)
    B = torch.from_numpy(b.copy()).cuda()
    handle.sytrs(T, perm, piv, B)
    handle.synchronize()
    x = B.cpu().numpy()
    assert np.allclose(A @ x, b, rtol=1e-8, atol=1e-8)


def test_synchronous_solve_reports_wait_error(oracle, handle, bbk):
    import torch
    n = 800
    A = oracle.gen_synth_symmetric(n, 4403, 1.0, 7)
    from blockedbunchkaufmanldlt_amd import padded_ld
    T = torch.zeros((n, padded_ld(n)), dtype=torch.float64, device="cuda")
    T[:, :n] = torch.from_numpy(np.triu(A)).cuda()
    b = torch.from_numpy(np.random.default_rng(2).standard_normal((n, 1))).cuda()
    X = torch.zeros_like(b)
    handle.set_wait_limit(-1)
    try:
        with pytest.raises(bbk.BbkError) as ei:
            handle.sysv(T, b, X)
        assert ei.value.code == bbk.BBK_ERR_HIP
    finally:
        handle.set_wait_limit(0)
    assert handle.sysv(T, b, X) < 1e-14


def test_info_reports_first_singular_pivot(handle):
    """The reference's fixed case diag(2, -3, 0, 5)
    (tests/block_ldlt_decomposition_test.cc:476-508): pivot 2 is exactly
    zero, so info = 3; a regular matrix gives 0."""
    A = np.diag([2.0, -3.0, 0.0, 5.0])
    _factor(handle, A, 2)
    assert handle.info() == (3, 0)
    B = np.array([[0.0, 1.0, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0],
                  [0.0, 0.0, 0.0, 1.0], [0.0, 0.0, 1.0, 0.0]])
    _factor(handle, B, 2)
    assert handle.info() == (0, 0)


def test_info_flags_nonfinite_input(oracle, handle):
    n = 300
    A = oracle.gen_synth_symmetric(n, 4404, 1.0, 0)
    A[5, 7] = A[7, 5] = np.nan
    T, perm, piv = _factor(handle, A)
    info, nonfinite = handle.info()
    assert nonfinite == 1
    # the factorization still ends normally: a valid permutation
    p = perm.cpu().numpy()
    assert np.array_equal(np.sort(p), np.arange(n))
    # a finite matrix afterwards clears the flag
    _factor(handle, oracle.gen_synth_symmetric(n, 4405, 1.0, 0))
    assert handle.info()[1] == 0

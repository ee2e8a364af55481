"""Out-of-core factorization (bbk_dsytrf_ooc): the matrix stays in pinned
host memory and HBM use for its working copies is bounded.  While the
trailing matrix is larger than the bound, the panel kernel runs on the host
copy across PCIe and the trailing update streams column slabs through HBM;
then the trailing matrix moves to HBM.  The kernels are the in-core ones and
every element takes the same arithmetic whichever slab holds it, so the
factors, perm and piv must be bitwise those of bbk_dsytrf (which the parity
tests pin to the oracle); the strictly lower triangle and the columns past n
must come back untouched.  Budgets are chosen to force each phase: the
smallest one the call accepts (one tile column per slab), budgets that
switch to HBM part way, and 0 (the whole matrix in HBM from the start)."""
import pytest

pytestmark = pytest.mark.gpu

LOW = -7.25   # strictly lower triangle sentinel
PAD = 3.5     # columns n .. lda-1


def _min_budget(n, nb):
    """The smallest device_bytes bbk_dsytrf_ooc accepts for (n, nb)."""
    def pitch(m):
        ld = (m + 63) // 64 * 64
        return ld + 64 if ld % 1024 == 0 else max(ld, 64)

    def r256(v):
        return (v + 255) // 256 * 256
    return r256(8 * (nb + 1) * pitch(n)) + 2 * r256(8 * n * 64)


def _compare(handle, n, nb, budget, seed, zero_stride=7, extra=3):
    import numpy as np
    import torch
    import blockedbunchkaufmanldlt_amd as B
    A = handle.gen_synth(n, seed, 1.0, zero_stride, B.padded_ld(n))
    T = A.clone()
    perm, piv = handle.sytrf(T, nb)
    handle.synchronize()
    lda = n + extra
    M = B.PinnedMatrix(n, lda)
    try:
        M.array[:, :n] = A[:, :n].cpu().numpy()
        low = np.tril_indices(n, -1)
        M.array[low] = LOW
        M.array[:, n:] = PAD
        p2, v2 = handle.sytrf_ooc(M, nb, budget)
        F = M.array.copy()
    finally:
        M.close()
    ref = np.triu(T[:, :n].cpu().numpy())
    got = np.triu(F[:, :n])
    assert np.array_equal(p2, perm.cpu().numpy()), (n, nb, budget)
    assert np.array_equal(v2, piv.cpu().numpy()), (n, nb, budget)
    diff = np.argwhere(got.view(np.uint64) != ref.view(np.uint64))
    assert diff.size == 0, (n, nb, budget, diff[:5])
    assert np.all(F[low] == LOW)
    assert np.all(F[:, n:] == PAD)
    return int((v2 == 2).sum())


@pytest.mark.parametrize("n,nb,budget", [
    (1000, 64, "min"),          # one tile column per slab until the end
    (1000, 64, 3 << 20),        # trailing matrix moves to HBM near m = 600
    (777, 32, 1 << 20),         # narrow panels, many straddling 2x2 pivots
    (2500, 128, 16 << 20),
    (300, 0, 0),                # whole matrix in HBM from the start
])
def test_ooc_is_bitwise_the_in_core_factorization(handle, n, nb, budget):
    if budget == "min":
        budget = _min_budget(n, nb if nb else 64)
    twos = _compare(handle, n, nb, budget, 9100 + n)
    assert twos > 0


def test_ooc_with_the_zero_diagonal_everywhere(handle):
    """Every diagonal entry zero: every pivot is a 2x2 or needs the candidate
    row, so every panel reads candidate rows across PCIe."""
    twos = _compare(handle, 900, 64, 2 << 20, 77, zero_stride=1)
    assert twos > 100


def test_ooc_larger_order(handle):
    """n = 6000 with a quarter of the matrix's bytes: about half the panels
    run on the host copy."""
    n = 6000
    budget = 8 * n * n // 4
    _compare(handle, n, 128, budget, 4242)


def test_ooc_refuses_unpinned_memory_and_small_budgets(handle):
    import ctypes
    import numpy as np
    import blockedbunchkaufmanldlt_amd as B
    n = 200
    A = np.zeros((n, n))
    perm = np.empty(n, dtype=np.int32)
    piv = np.empty(n, dtype=np.int32)
    rc = B.lib().bbk_dsytrf_ooc(handle.h, n, 32, A.ctypes.data, n, perm.ctypes.data,
                                piv.ctypes.data, 0)
    assert rc == -4
    assert b"pinned" in B.lib().bbk_last_error(handle.h)
    M = B.PinnedMatrix(n)
    try:
        M.array[:] = np.eye(n)
        with pytest.raises(B.BbkError) as err:
            handle.sytrf_ooc(M, 32, _min_budget(n, 32) - 1)
        assert err.value.code == B.BBK_ERR_UNSUPPORTED
        # a leading dimension that runs past the allocation is refused
        rc = B.lib().bbk_dsytrf_ooc(handle.h, n, 32, M.ptr, n + 1, perm.ctypes.data,
                                    piv.ctypes.data, 0)
        assert rc == -4
        # the accepted minimum factors the identity
        p, v = handle.sytrf_ooc(M, 32, _min_budget(n, 32))
        assert np.array_equal(p, np.arange(n))
        assert np.all(v == 1)
        assert np.array_equal(np.triu(M.array), np.eye(n))
    finally:
        M.close()
    del ctypes


def test_ooc_with_the_wide_panel_variant(handle):
    """The 512-column panel workgroups (the variant orders past 65536 use),
    forced on both paths: the out-of-core factors stay bitwise the in-core
    ones."""
    handle.set_panel_threads(512)
    try:
        twos = _compare(handle, 3000, 64, 8 * 3000 * 3000 // 3, 6464)
    finally:
        handle.set_panel_threads(0)
    assert twos > 0

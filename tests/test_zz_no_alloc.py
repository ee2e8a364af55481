"""GPU test of bbk_dsytrf_noalloc (the factorization that allocates
nothing, its scratch in a fixed device array), through the C ABI.

Bar: bitwise the same compact factors, permutation and pivot types as
bbk_dsytrf on the same input (same kernels), on consecutive calls whose
orders shrink (the largest supported order first), so a later call runs on
the array the earlier one left behind; residual < 1e-10 on the device.  The
host-side checks reject an order above the array's limit and the null
stream before any device work.
"""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _entry(bbk):
    L = bbk.lib()
    fn = L.bbk_dsytrf_noalloc
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64,
                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    return fn


def test_noalloc_matches_dsytrf_on_shrinking_calls(bbk, handle):
    import torch
    fn = _entry(bbk)
    stream = handle.stream.cuda_stream
    assert stream  # the handle's own non-blocking stream
    for n, nb, seed in ((16384, 128, 31), (3000, 128, 32), (1000, 64, 33)):
        lda = bbk.padded_ld(n)
        A = handle.gen_synth(n, seed, 1.0, 7, lda=lda)
        handle.synchronize()
        ref = A.clone()
        mine = A.clone()
        p_ref, v_ref = handle.sytrf(ref, nb)
        handle.synchronize()
        perm = torch.full((n,), -1, dtype=torch.int32, device="cuda")
        piv = torch.full((n,), -9, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        print(f"noalloc n={n} nb={nb} seed={seed}", flush=True)
        rc = fn(n, nb, mine.data_ptr(), lda, perm.data_ptr(), piv.data_ptr(), stream)
        assert rc == 0, rc
        torch.cuda.synchronize()
        assert torch.equal(torch.triu(mine[:, :n]), torch.triu(ref[:, :n])), n
        assert torch.equal(perm, p_ref) and torch.equal(piv, v_ref), n
        if n <= 3000:
            assert handle.residual(A, mine, perm, piv) < 1e-10
        del A, ref, mine


def test_noalloc_argument_checks(bbk, handle):
    fn = _entry(bbk)
    stream = handle.stream.cuda_stream
    # no device work happens on these paths
    assert fn(16385, 128, 1, 16448, 1, 1, stream) == bbk.BBK_ERR_UNSUPPORTED
    assert fn(100, 32, 1, 128, 1, 1, None) == -7
    assert fn(0, 32, 1, 128, 1, 1, stream) == -1
    assert fn(100, 32, 1, 99, 1, 1, stream) == -4

"""The largest supported order.  The panel kernel keeps every workgroup of
its grid co-resident, one per CU at most 256 of them, each holding 256
columns, so n = 65536 is the largest matrix (34 GB in binary64; row offsets
pass 2^31 elements, so every leading-dimension product must be 64-bit).
One order more is refused with BBK_ERR_UNSUPPORTED before any launch."""
import pytest

pytestmark = pytest.mark.gpu


def test_largest_order_factors(handle):
    import torch
    import blockedbunchkaufmanldlt_amd as B
    n = 65536
    lda = B.padded_ld(n)
    A = handle.gen_synth(n, 91, 1.0, 7, lda)
    T = A.clone()
    perm, piv = handle.sytrf(T)
    rs = handle.sampled_residual(A, T, perm, piv, 3)
    handle.synchronize()
    assert rs < 1e-10, rs
    assert torch.equal(torch.sort(perm.long())[0], torch.arange(n, device=perm.device))
    assert int((piv == 2).sum()) > 0
    del A, T
    torch.cuda.empty_cache()


def test_order_past_the_panel_grid_is_refused(handle):
    import torch
    import blockedbunchkaufmanldlt_amd as B
    n = 65537
    # a buffer of the full size, so that nothing could be out of bounds even
    # if the size check did not hold
    T = torch.zeros((n, B.padded_ld(n)), dtype=torch.float64, device="cuda")
    with pytest.raises(B.BbkError) as err:
        handle.sytrf(T)
    assert err.value.code == B.BBK_ERR_UNSUPPORTED
    del T
    torch.cuda.empty_cache()

"""The largest supported orders.  The panel kernel keeps every workgroup of
its grid co-resident, at most 256 of them (one per CU).  Workgroups of 256
columns carry orders up to 65536; past that the 512-column variant (binary64
panels at most 64 wide) carries orders up to 131072 (137 GB in binary64).
Row offsets pass 2^31 elements, so every leading-dimension product must be
64-bit.  One order more than the largest is refused with
BBK_ERR_UNSUPPORTED before any launch.

The two variants make the same pivot decisions and, at the same panel
width, the same arithmetic in the same order, so their factors are bitwise
equal: that pins the 512-column variant to the 256-column one, which the
parity tests pin to the oracle."""
import pytest

pytestmark = pytest.mark.gpu


def _factor_both(handle, n, dtype, nb, seed):
    import torch
    import blockedbunchkaufmanldlt_amd as B
    lda = B.padded_ld(n)
    A = handle.gen_synth(n, seed, 1.0, 7, lda)
    if dtype == torch.float32:
        A = A.to(torch.float32)
    out = []
    try:
        for nt in (256, 512):
            handle.set_panel_threads(nt)
            T = A.clone()
            perm, piv = handle.sytrf(T, nb)
            handle.synchronize()
            out.append((T, perm, piv))
    finally:
        handle.set_panel_threads(0)
    return A, out


@pytest.mark.parametrize("n,dtype_name,nb", [(3000, "f64", 64), (3001, "f32", 128),
                                              (20000, "f64", 64), (600, "f64", 32)])
def test_wide_panel_variant_is_bitwise_the_standard_one(handle, n, dtype_name, nb):
    import torch
    dtype = torch.float64 if dtype_name == "f64" else torch.float32
    A, ((T1, p1, v1), (T2, p2, v2)) = _factor_both(handle, n, dtype, nb, 7100 + n)
    assert torch.equal(p1, p2)
    assert torch.equal(v1, v2)
    assert torch.equal(torch.triu(T1[:, :n]), torch.triu(T2[:, :n]))
    assert int((v1 == 2).sum()) > 0
    del A, T1, T2
    torch.cuda.empty_cache()


def test_forced_variant_refuses_an_order_its_grid_cannot_hold(handle):
    import torch
    import blockedbunchkaufmanldlt_amd as B
    n = 65537
    T = torch.zeros((n, B.padded_ld(n)), dtype=torch.float64, device="cuda")
    handle.set_panel_threads(256)
    try:
        with pytest.raises(B.BbkError) as err:
            handle.sytrf(T)
        assert err.value.code == B.BBK_ERR_UNSUPPORTED
    finally:
        handle.set_panel_threads(0)
    del T
    torch.cuda.empty_cache()


def test_largest_standard_order_factors(handle):
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


def test_order_past_65536_factors(handle):
    """n = 98304 (77 GB per copy in binary64) through the 512-column panel
    variant, checked by the reference stress test's sampled residual."""
    import torch
    import blockedbunchkaufmanldlt_amd as B
    n = 98304
    lda = B.padded_ld(n)
    A = handle.gen_synth(n, 92, 1.0, 7, lda)
    T = A.clone()
    perm, piv = handle.sytrf(T)
    rs = handle.sampled_residual(A, T, perm, piv, 5)
    handle.synchronize()
    assert rs < 1e-10, rs
    assert torch.equal(torch.sort(perm.long())[0], torch.arange(n, device=perm.device))
    assert int((piv == 2).sum()) > 0
    assert handle.info() == (0, 0)
    del A, T
    torch.cuda.empty_cache()


def test_order_past_the_panel_grid_is_refused(handle):
    import torch
    import blockedbunchkaufmanldlt_amd as B
    n = 131073
    # a buffer of the full size (137 GB), so that nothing could be out of
    # bounds even if the size check did not hold
    T = torch.zeros((n, B.padded_ld(n)), dtype=torch.float64, device="cuda")
    with pytest.raises(B.BbkError) as err:
        handle.sytrf(T)
    assert err.value.code == B.BBK_ERR_UNSUPPORTED
    del T
    torch.cuda.empty_cache()

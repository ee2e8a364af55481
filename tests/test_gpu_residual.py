"""GPU tests of the reconstruction residuals (bbk_dresidual,
bbk_dsampled_residual) against the oracle's restatements of the reference's
RelativeError(ReconstructOriginal) and SampledRelativeResidual
(tests/block_ldlt_decomposition_test.cc:131-135,
tests/block_ldlt_large_stress_test.cc:160-244).

On a correct factorization both residuals sit at rounding level, where the
GPU's summation order and the oracle's differ by as much as the values
themselves; so each case also plants an error in the compact factor (an L
entry, a D entry, a 2x2 pivot's off-diagonal D entry) and requires the two
sides to agree to 1e-9 on the perturbed factor, where rounding is
negligible.  A wrong index, sign, pivot-type case or sample position changes
the perturbed residual by far more than that."""
import numpy as np
import pytest

from gpu_util import to_device

pytestmark = pytest.mark.gpu


def factor(handle, A, nb=0):
    import torch
    n = A.shape[0]
    T = to_device(A)
    perm, piv = handle.sytrf(T, nb)
    handle.synchronize()
    return T, perm, piv


def planted(F, piv, n):
    """(row, col) entries of the upper-triangle factor to perturb: an L
    entry, a 1x1 D entry and, when there is one, a 2x2 pivot's D(k, k+1)."""
    pv = piv.cpu().numpy()
    out = [(3, n - 2), (n // 2, n // 2)]
    ks = np.nonzero(pv == 2)[0]
    if len(ks):
        k = int(ks[len(ks) // 2])
        out.append((k, k + 1))
    return out


@pytest.mark.parametrize("n,nb", [(300, 32), (513, 64), (1000, 128)])
def test_full_residual_matches_oracle(oracle, handle, n, nb):
    A = oracle.gen_synth_symmetric(n, 700 + n)
    T, perm, piv = factor(handle, A, nb)
    Ad = to_device(A)
    F = np.triu(T[:, :n].cpu().numpy())
    g = oracle.Factorization(n, nb, F, perm.cpu().numpy().astype(np.int32),
                             piv.cpu().numpy().astype(np.int32))
    r_gpu = handle.residual(Ad, T, perm, piv)
    r_cpu = oracle.relative_residual(A, g)
    assert r_gpu < 1e-13 and r_cpu < 1e-13
    assert abs(r_gpu - r_cpu) < 1e-14
    for (i, j) in planted(F, piv, n):
        Tp = T.clone()
        Tp[i, j] += 1e-3 * (1.0 + abs(float(Tp[i, j])))
        Fp = np.triu(Tp[:, :n].cpu().numpy())
        gp = oracle.Factorization(n, nb, Fp, g.perm, g.piv)
        rg = handle.residual(Ad, Tp, perm, piv)
        rc = oracle.relative_residual(A, gp)
        assert rc > 1e-7, (i, j, rc)
        assert abs(rg - rc) <= 1e-9 * rc, (i, j, rg, rc)


@pytest.mark.parametrize("n,seed", [(300, 5), (777, 11)])
def test_sampled_residual_matches_oracle(oracle, handle, n, seed):
    A = oracle.gen_synth_symmetric(n, 900 + n, 1e3)
    T, perm, piv = factor(handle, A, 64)
    Ad = to_device(A)
    F = np.triu(T[:, :n].cpu().numpy())
    g = oracle.Factorization(n, 64, F, perm.cpu().numpy().astype(np.int32),
                             piv.cpu().numpy().astype(np.int32))
    r_gpu = handle.sampled_residual(Ad, T, perm, piv, seed)
    r_cpu = oracle.sampled_residual(A, g, seed)
    assert r_gpu < 1e-12 and r_cpu < 1e-12
    assert abs(r_gpu - r_cpu) < 1e-14
    # a planted L entry in the row of a sampled position: the perturbed
    # residual depends on exactly which samples hit it
    rows, cols = oracle.residual_samples(n, seed)
    inv = np.empty(n, dtype=np.int64)
    inv[g.perm] = np.arange(n)
    fr = int(inv[rows[10]])
    targets = [(0, fr)] if fr > 0 else []
    targets += planted(F, piv, n)
    for (i, j) in targets:
        Tp = T.clone()
        Tp[i, j] += 1e-2 * (1.0 + abs(float(Tp[i, j])))
        Fp = np.triu(Tp[:, :n].cpu().numpy())
        gp = oracle.Factorization(n, 64, Fp, g.perm, g.piv)
        rg = handle.sampled_residual(Ad, Tp, perm, piv, seed)
        rc = oracle.sampled_residual(A, gp, seed)
        if rc > 1e-9:
            assert abs(rg - rc) <= 1e-9 * rc, (i, j, rg, rc)
        else:  # no sample touches the entry
            assert rg < 1e-12


def test_residuals_at_full_size(handle):
    """n = 16384 (the benchmark order): the whole reconstruction residual
    and the stress test's sampled one, on the device, at rounding level."""
    import torch
    import blockedbunchkaufmanldlt_amd as B
    n = 16384
    lda = B.padded_ld(n)
    A = handle.gen_synth(n, 31, 1.0, 0, lda)
    T = A.clone()
    perm, piv = handle.sytrf(T)
    r = handle.residual(A, T, perm, piv)
    rs = handle.sampled_residual(A, T, perm, piv, 7)
    handle.synchronize()
    assert r < 1e-12, r
    assert rs < 1e-12, rs
    assert int((piv == 2).sum()) > 0


def test_residual_argument_errors(handle):
    import blockedbunchkaufmanldlt_amd as B
    import ctypes
    out = ctypes.c_double()
    L = B.lib()
    assert L.bbk_dresidual(handle.h, 0, None, 1, None, 1, None, None, ctypes.byref(out)) == -2
    assert L.bbk_dsampled_residual(handle.h, 4, 1, 4, 1, 4, 1, 1, 0, 2,
                                   ctypes.byref(out)) == -11


@pytest.mark.parametrize("n", [1, 5, 130, 400])
def test_reconstruct_matches_dense_product(oracle, handle, n):
    """bbk_dreconstruct against P L D L^T P^T built densely from the
    oracle's ExtractRegularFactorsFromCompact restatement (same factors)."""
    A = oracle.gen_synth_symmetric(n, 60 + n)
    T, perm, piv = factor(handle, A, 32)
    F = np.triu(T[:, :n].cpu().numpy())
    g = oracle.Factorization(n, 32, F, perm.cpu().numpy().astype(np.int32),
                             piv.cpu().numpy().astype(np.int32))
    L, D = oracle.extract_ld(g)
    M = L @ D @ L.T
    R = np.empty_like(M)
    R[np.ix_(g.perm, g.perm)] = M
    out = handle.reconstruct(T, perm, piv).cpu().numpy()
    assert np.allclose(out, out.T, rtol=0, atol=0)
    scale = max(1.0, np.abs(M).max())
    assert np.abs(out - R).max() <= 1e-13 * scale * max(1, n)
    assert np.abs(out - A).max() <= 1e-12 * scale * max(1, n)


def test_fp32_factorization_residual_at_full_size(handle):
    """The binary32 form at n = 16384: its factor, widened, reconstructs the
    binary64 input to binary32 accuracy (checked on the device)."""
    import torch
    import blockedbunchkaufmanldlt_amd as B
    n = 16384
    lda = B.padded_ld(n)
    A = handle.gen_synth(n, 41, 1.0, 0, lda)
    T32 = A.to(torch.float32)
    perm, piv = handle.sytrf(T32)
    handle.synchronize()
    F = T32.to(torch.float64)
    del T32
    r = handle.residual(A, F, perm, piv)
    assert 1e-9 < r < 1e-4, r


def test_factor_and_residual_at_double_size(handle):
    """n = 32768 (8.6 GB in binary64, 128 panel workgroups): the device
    reconstruction residual of the factor, and a valid permutation."""
    import torch
    import blockedbunchkaufmanldlt_amd as B
    n = 32768
    lda = B.padded_ld(n)
    A = handle.gen_synth(n, 77, 1.0, 7, lda)
    T = A.clone()
    perm, piv = handle.sytrf(T)
    r = handle.residual(A, T, perm, piv)
    handle.synchronize()
    assert r < 1e-11, r
    assert torch.equal(torch.sort(perm.long())[0], torch.arange(n, device=perm.device))
    del A, T
    torch.cuda.empty_cache()

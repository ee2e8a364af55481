"""Helpers shared by the GPU tests: move numpy matrices through the C ABI."""
import numpy as np


def to_device(A, dtype="float64", lda=None):
    import torch
    import blockedbunchkaufmanldlt_amd as B
    n = A.shape[0]
    lda = lda or B.padded_ld(n)
    tdt = torch.float64 if dtype == "float64" else torch.float32
    T = torch.zeros((n, lda), dtype=tdt, device="cuda")
    T[:, :n] = torch.from_numpy(np.ascontiguousarray(np.triu(A))).to(tdt)
    return T


def gpu_factor(handle, oracle, A, nb=0, dtype="float64", lda=None):
    """Factor A's upper triangle on the GPU; returns an oracle.Factorization
    holding the upper triangle of the device result (as float64)."""
    n = A.shape[0]
    T = to_device(A, dtype, lda)
    perm, piv = handle.sytrf(T, nb)
    handle.synchronize()
    F = np.triu(T[:, :n].double().cpu().numpy())
    return oracle.Factorization(n, nb, F, perm.cpu().numpy().astype(np.int32),
                                piv.cpu().numpy().astype(np.int32))

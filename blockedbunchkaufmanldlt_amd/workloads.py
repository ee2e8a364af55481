"""Synthetic inputs of the five benchmark workloads, built on the GPU.

C1  n = 512, panel width 64: A(i, j) = A(j, i) uniform in [-1, 1) from the
    SplitMix64 generator keyed by (seed, i, j), diagonal included.
C2  n = 5000 / 10000, panel width 128: the same entries times a scale in
    {1, 1e-6, 1e6}, zero diagonal at every index i % 7 == 0.
C3  n = 16384: A = Q diag(lambda) Q^T, Q from the QR of a seeded Gaussian
    matrix, |lambda| log-spaced over [1e-6, 1] with seeded random signs
    (condition 1e6); b ~ N(0, 1).
C4  4096 matrices, n uniform in [16, 256], uniform [-1, 1) entries.
C5  1024 matrices, n uniform in [256, 2048], condition 10^U[4, 7] per
    matrix: A = Q diag(lambda) Q^T with Q from the QR of a seeded Gaussian
    matrix (a dense random orthogonal matrix), |lambda| log-spaced over
    [1/kappa, 1] with random signs; four right-hand sides ~ N(0, 1).

Used by bench.py and the GPU tests, so both see the same matrices.
"""
from __future__ import annotations

import math
from typing import List, Tuple


def c3_matrix(torch, n: int, seed: int, lda: int, kappa: float = 1e6):
    """(A, b): A = Q diag(lambda) Q^T in an (n, lda) float64 tensor, Q from
    the QR of a seeded Gaussian matrix (on the GPU); b ~ N(0, 1), (n, 1)."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    G = torch.randn((n, n), generator=gen, dtype=torch.float64, device="cuda")
    Q, _ = torch.linalg.qr(G)
    del G
    i = torch.arange(n, dtype=torch.float64, device="cuda")
    lam = torch.pow(torch.tensor(10.0, dtype=torch.float64, device="cuda"),
                    -math.log10(kappa) * i / max(n - 1, 1))
    sgn = torch.randint(0, 2, (n,), generator=gen, device="cuda").to(torch.float64) * 2 - 1
    A = (Q * (lam * sgn)) @ Q.T.contiguous()
    del Q
    A = 0.5 * (A + A.T)
    out = torch.zeros((n, lda), dtype=torch.float64, device="cuda")
    out[:, :n] = A
    del A
    b = torch.randn((n, 1), generator=gen, dtype=torch.float64, device="cuda")
    return out, b


def householder_eig(torch, n: int, kappa: float, gen, reflectors: int = 4):
    """A = U diag(lambda) U^T with U a product of a few random Householder
    reflectors: exact eigenvalues |lambda| log-spaced over [1/kappa, 1] with
    random signs, but A is diag(lambda) plus a rank-2 * reflectors term, a
    cheap, strongly structured test matrix (small tests only)."""
    lam = torch.pow(torch.tensor(10.0, dtype=torch.float64, device="cuda"),
                    -math.log10(kappa) * torch.arange(n, dtype=torch.float64, device="cuda")
                    / max(n - 1, 1))
    sgn = torch.randint(0, 2, (n,), generator=gen, device="cuda").to(torch.float64) * 2 - 1
    A = torch.diag(lam * sgn)
    for _ in range(reflectors):
        v = torch.randn(n, generator=gen, dtype=torch.float64, device="cuda")
        v = v / torch.linalg.norm(v)
        Av = A @ v
        A = A - 2 * torch.outer(v, Av) - 2 * torch.outer(Av, v) + \
            4 * torch.dot(v, Av) * torch.outer(v, v)
    return 0.5 * (A + A.T)


def orthogonal_eig(torch, n: int, kappa: float, gen):
    """A = Q diag(lambda) Q^T with Q from the QR of a seeded N(0, 1) matrix
    (a dense random orthogonal matrix, binary64 on the GPU): eigenvalues
    |lambda| log-spaced over [1/kappa, 1] with random signs, so the
    condition number is kappa up to Q's orthogonality error."""
    G = torch.randn((n, n), generator=gen, dtype=torch.float64, device="cuda")
    Q, _ = torch.linalg.qr(G)
    del G
    lam = torch.pow(torch.tensor(10.0, dtype=torch.float64, device="cuda"),
                    -math.log10(kappa) * torch.arange(n, dtype=torch.float64, device="cuda")
                    / max(n - 1, 1))
    sgn = torch.randint(0, 2, (n,), generator=gen, device="cuda").to(torch.float64) * 2 - 1
    A = (Q * (lam * sgn)) @ Q.T.contiguous()
    return 0.5 * (A + A.T)


def c4_sizes(seed: int = 4096, count: int = 4096) -> List[int]:
    import numpy as np
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.integers(16, 257, size=count)]


def c5_spec(seed: int = 1024, count: int = 1024) -> Tuple[List[int], List[float]]:
    """Orders and condition numbers of the batched mixed workload."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ns = [int(x) for x in rng.integers(256, 2049, size=count)]
    kap = [float(k) for k in 10.0 ** rng.uniform(4, 7, size=count)]
    return ns, kap


def c5_batch(torch, padded_ld, seed: int = 1024, count: int = 1024, nrhs: int = 4):
    """(ns, mats, Bs): the C5 matrices as (n, padded_ld(n)) float64 tensors
    (full symmetric storage) and their right-hand sides (n, nrhs)."""
    ns, kap = c5_spec(seed, count)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    mats = []
    for n, k in zip(ns, kap):
        T = torch.empty((n, padded_ld(n)), dtype=torch.float64, device="cuda")
        T[:, :n] = orthogonal_eig(torch, n, k, gen)
        mats.append(T)
    Bs = [torch.randn((n, nrhs), generator=gen, dtype=torch.float64, device="cuda") for n in ns]
    return ns, mats, Bs

"""ctypes bindings for the oracle (TEST INFRASTRUCTURE, never the product).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py import this package.  It wraps:

* ``liboracle.so`` -- the C restatement of the reference's blocked
  Bunch-Kaufman factorization (bk_oracle.c) plus libstdc++-compatible input
  generators (bk_gen.cc);
* ``_ref/libbkref.so`` -- the reference itself, compiled from its own
  sources by oracle/Makefile, when it has been built.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libbkref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Decision(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("r", ctypes.c_int),
                ("omega", ctypes.c_double), ("omega2", ctypes.c_double),
                ("akk", ctypes.c_double), ("sigma", ctypes.c_double),
                ("arr", ctypes.c_double)]


def build(with_reference: bool = True) -> None:
    target = "all" if with_reference else "oracle"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build(with_reference=False)
        L = ctypes.CDLL(ORACLE_SO)
        L.bko_dsytrf.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _ip, _ip,
                                 _ip, _ip, ctypes.POINTER(Decision)]
        L.bko_dsytrf_omp.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _ip,
                                     _ip, _ip, _ip, ctypes.POINTER(Decision)]
        L.bko_relative_residual.argtypes = [ctypes.c_int, _dp, _dp, _ip, _ip]
        L.bko_relative_residual.restype = ctypes.c_double
        L.bko_sampled_residual.argtypes = [ctypes.c_int, _dp, _dp, _ip, _ip,
                                           ctypes.c_int, _ip, _ip]
        L.bko_sampled_residual.restype = ctypes.c_double
        L.bko_reconstructed_entry.argtypes = [ctypes.c_int, _dp, _ip,
                                              ctypes.c_int, ctypes.c_int]
        L.bko_reconstructed_entry.restype = ctypes.c_double
        L.bko_extract_ld.argtypes = [ctypes.c_int, _dp, _ip, _dp, _dp]
        L.bko_dsytrs.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _ip, _ip, _dp]
        L.bko_gen_reference_symmetric.argtypes = [ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_double, _dp]
        L.bko_gen_stress_matrix.argtypes = [ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_double, _dp]
        L.bko_residual_samples.argtypes = [ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, _ip, _ip]
        L.bko_synth_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int,
                                        ctypes.c_int]
        L.bko_synth_uniform.restype = ctypes.c_double
        L.bko_gen_synth_symmetric.argtypes = [ctypes.c_int, ctypes.c_uint64,
                                              ctypes.c_double, ctypes.c_int, _dp]
        L.bko_kernel_scale_1x1.argtypes = [ctypes.c_int, ctypes.c_double, _dp,
                                           ctypes.c_int, _dp, ctypes.c_int]
        L.bko_kernel_scale_2x2.argtypes = [ctypes.c_int] + [ctypes.c_double] * 3 + \
            [_dp, ctypes.c_int] * 4
        L.bko_kernel_update_trailing.argtypes = [ctypes.c_int] * 3 + \
            [_dp, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(REF_SO)
        R = ctypes.CDLL(REF_SO)
        R.ref_factorize.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                    ctypes.c_int, _dp, _ip, _ip, _ip, _dp, _dp]
        R.ref_factorize.restype = ctypes.c_int
        R.ref_time_factorize.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp,
                                         _dp, _ip, _ip]
        R.ref_time_factorize.restype = ctypes.c_int
        R.ref_time_batch.argtypes = [ctypes.c_int, _ip, _dp, ctypes.POINTER(ctypes.c_longlong),
                                     ctypes.c_int, _dp]
        R.ref_time_batch.restype = ctypes.c_int
        R.ref_kernel_scale_1x1.argtypes = [ctypes.c_int, ctypes.c_double, _dp,
                                           ctypes.c_int, _dp, ctypes.c_int]
        R.ref_kernel_scale_2x2.argtypes = [ctypes.c_int] + [ctypes.c_double] * 3 + \
            [_dp, ctypes.c_int] * 4
        R.ref_kernel_update_trailing.argtypes = [ctypes.c_int] * 3 + \
            [_dp, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int]
        # helpers of the reference's own test programs (oracle/ref_tests_*.cc)
        R.ref_random_symmetric_matrix.argtypes = [ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_double, _dp]
        R.ref_random_symmetric_block_matrix.argtypes = [ctypes.c_int, ctypes.c_int,
                                                        ctypes.c_int, ctypes.c_double, _dp]
        R.ref_residual_samples.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip]
        R.ref_residual_samples.restype = ctypes.c_int
        R.ref_stress_case.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, _dp, _ip, _ip, _dp]
        R.ref_stress_case.restype = ctypes.c_int
        _ref = R
    return _ref


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


class Factorization:
    """Compact factors + permutation + pivot types, reference encoding."""

    def __init__(self, n, bs, compact, perm, piv, panel_start=None,
                 panel_end=None, trace=None, workspace=None, geometry=None,
                 residual=None):
        self.n, self.bs = n, bs
        self.compact = compact          # n x n, upper triangle meaningful
        self.perm = perm
        self.piv = piv
        self.panel_start = panel_start
        self.panel_end = panel_end
        self.trace = trace
        self.workspace = workspace
        self.geometry = geometry
        self.residual = residual


def dsytrf(A: np.ndarray, bs: int, trace: bool = False, omp: Optional[bool] = None) -> Factorization:
    """Oracle factorization of the symmetric matrix whose upper triangle is A's."""
    n = A.shape[0]
    F = np.ascontiguousarray(np.triu(A), dtype=np.float64)
    W = np.zeros((n, n), dtype=np.float64)
    perm = np.zeros(n, dtype=np.int32)
    piv = np.zeros(n, dtype=np.int32)
    nb = (n + bs - 1) // bs
    ps = np.zeros(nb, dtype=np.int32)
    pe = np.zeros(nb, dtype=np.int32)
    tr = (Decision * n)() if trace else None
    if omp is None:
        omp = n >= 400
    if omp:
        rc = lib().bko_dsytrf_omp(n, bs, _d(F), _d(W), _i(perm), _i(piv), _i(ps), _i(pe),
                                  tr if trace else None)
    else:
        rc = lib().bko_dsytrf(n, bs, _d(F), _d(W), _i(perm), _i(piv), _i(ps), _i(pe),
                              tr if trace else None)
    if rc != 0:
        raise ValueError("oracle rejected arguments")
    return Factorization(n, bs, np.triu(F), perm, piv, ps, pe,
                         list(tr) if trace else None, np.triu(W))


def ref_dsytrf(A: np.ndarray, bs: int, full: bool = True, lower: bool = False) -> Factorization:
    """The reference itself (oracle/_ref), as its tests drive it."""
    n = A.shape[0]
    A = np.ascontiguousarray(A, dtype=np.float64)
    compact = np.zeros((n, n))
    work = np.zeros((n, n))
    perm = np.zeros(n, dtype=np.int32)
    piv = np.zeros(n, dtype=np.int32)
    nb = (n + bs - 1) // bs
    geom = np.zeros(5 * nb * (nb + 1) // 2, dtype=np.int32)
    res = np.zeros(1)
    rc = ref().ref_factorize(n, bs, _d(A), int(lower), int(full), _d(compact), _i(perm),
                             _i(piv), _i(geom), _d(work), _d(res))
    if rc != 0:
        raise RuntimeError("reference threw")
    return Factorization(n, bs, compact, perm, piv, workspace=work,
                         geometry=geom.reshape(-1, 5), residual=float(res[0]) if full else None)


def ref_time_factorize(A: np.ndarray, bs: int, want_factors: bool = False):
    n = A.shape[0]
    A = np.ascontiguousarray(A, dtype=np.float64)
    sec = np.zeros(1)
    compact = np.zeros((n, n)) if want_factors else None
    perm = np.zeros(n, dtype=np.int32) if want_factors else None
    piv = np.zeros(n, dtype=np.int32) if want_factors else None
    rc = ref().ref_time_factorize(n, bs, _d(A), _d(sec),
                                  _d(compact) if want_factors else None,
                                  _i(perm) if want_factors else None,
                                  _i(piv) if want_factors else None)
    if rc != 0:
        raise RuntimeError("reference threw")
    if want_factors:
        return float(sec[0]), Factorization(n, bs, compact, perm, piv)
    return float(sec[0])


def ref_time_batch(mats, bs: int) -> float:
    """Seconds of the reference's Factorize() over a list of symmetric
    matrices, an outer OpenMP loop over the matrices with each factorization
    on one thread (batched CPU baseline)."""
    ns = np.array([m.shape[0] for m in mats], dtype=np.int32)
    offs = np.zeros(len(mats), dtype=np.int64)
    total = 0
    for i, m in enumerate(mats):
        offs[i] = total
        total += m.shape[0] * m.shape[0]
    data = np.empty(max(total, 1), dtype=np.float64)
    for m, o in zip(mats, offs):
        data[o:o + m.size] = np.ascontiguousarray(m, dtype=np.float64).ravel()
    sec = np.zeros(1)
    fails = ref().ref_time_batch(len(mats), _i(ns), _d(data),
                                 offs.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), bs,
                                 _d(sec))
    if fails:
        raise RuntimeError(f"reference threw on {fails} matrices")
    return float(sec[0])


# The reference's own test helpers (checker library only).
REF_SAMPLE_COUNT = 4096  # kResidualSampleCount, tests/block_ldlt_large_stress_test.cc:25


def ref_random_symmetric_matrix(n: int, seed: int, scale: float = 1.0) -> np.ndarray:
    out = np.zeros((n, n))
    ref().ref_random_symmetric_matrix(n, seed, scale, _d(out))
    return out


def ref_random_symmetric_block_matrix(n: int, bs: int, seed: int, scale: float = 1.0) -> np.ndarray:
    out = np.zeros((n, n))
    ref().ref_random_symmetric_block_matrix(n, bs, seed, scale, _d(out))
    return out


def ref_residual_samples(n: int, seed: int):
    rows = np.zeros(REF_SAMPLE_COUNT, dtype=np.int32)
    cols = np.zeros(REF_SAMPLE_COUNT, dtype=np.int32)
    cnt = ref().ref_residual_samples(n, seed, _i(rows), _i(cols))
    return rows[:cnt], cols[:cnt]


def ref_stress_case(n: int, bs: int, seed: int, scale: float = 1.0):
    """The large stress test's case: its generator, a compact-only
    Factorize() and its SampledRelativeResidual.  Returns
    (Factorization, sampled residual)."""
    compact = np.zeros((n, n))
    perm = np.zeros(n, dtype=np.int32)
    piv = np.zeros(n, dtype=np.int32)
    res = np.zeros(1)
    if ref().ref_stress_case(n, bs, seed, scale, _d(compact), _i(perm), _i(piv), _d(res)) != 0:
        raise RuntimeError("reference threw")
    return Factorization(n, bs, compact, perm, piv), float(res[0])


def relative_residual(A: np.ndarray, f: Factorization) -> float:
    A = np.ascontiguousarray(A, dtype=np.float64)
    F = np.ascontiguousarray(f.compact, dtype=np.float64)
    return lib().bko_relative_residual(A.shape[0], _d(A), _d(F), _i(f.perm), _i(f.piv))


def residual_samples(n: int, seed: int, count: int = 4096):
    rows = np.zeros(count, dtype=np.int32)
    cols = np.zeros(count, dtype=np.int32)
    lib().bko_residual_samples(n, seed, count, _i(rows), _i(cols))
    return rows, cols


def sampled_residual(A: np.ndarray, f: Factorization, seed: int, count: int = 4096) -> float:
    rows, cols = residual_samples(A.shape[0], seed, count)
    A = np.ascontiguousarray(A, dtype=np.float64)
    F = np.ascontiguousarray(f.compact, dtype=np.float64)
    return lib().bko_sampled_residual(A.shape[0], _d(A), _d(F), _i(f.perm), _i(f.piv),
                                      count, _i(rows), _i(cols))


def extract_ld(f: Factorization):
    n = f.n
    L = np.zeros((n, n))
    D = np.zeros((n, n))
    F = np.ascontiguousarray(f.compact, dtype=np.float64)
    lib().bko_extract_ld(n, _d(F), _i(f.piv), _d(L), _d(D))
    return L, D


def dsytrs(f: Factorization, b: np.ndarray) -> np.ndarray:
    b2 = np.array(b, dtype=np.float64, order="C", copy=True)
    nrhs = 1 if b2.ndim == 1 else b2.shape[1]
    F = np.ascontiguousarray(f.compact, dtype=np.float64)
    lib().bko_dsytrs(f.n, nrhs, _d(F), _i(f.perm), _i(f.piv), _d(b2))
    return b2


def gen_reference_symmetric(n: int, seed: int, scale: float = 1.0) -> np.ndarray:
    out = np.zeros((n, n))
    lib().bko_gen_reference_symmetric(n, seed, scale, _d(out))
    return out


def gen_stress_matrix(n: int, bs: int, seed: int, scale: float = 1.0) -> np.ndarray:
    out = np.zeros((n, n))
    lib().bko_gen_stress_matrix(n, bs, seed, scale, _d(out))
    return out


def gen_synth_symmetric(n: int, seed: int, scale: float = 1.0, zero_diag_stride: int = 0) -> np.ndarray:
    out = np.zeros((n, n))
    lib().bko_gen_synth_symmetric(n, seed, scale, zero_diag_stride, _d(out))
    return out


def synth_uniform(seed: int, i: int, j: int) -> float:
    return lib().bko_synth_uniform(seed, i, j)


def lapack_to_reference(ipiv: np.ndarray):
    """Map LAPACK dsytrf('L') ipiv (1-based) to (perm, piv) in the
    reference's encoding, following the forward-order, by-factor-position
    interpretation of tests/block_ldlt_decomposition_test.cc:196-240."""
    n = len(ipiv)
    perm = list(range(n))
    piv = [0] * n
    k = 0
    while k < n:
        if ipiv[k] > 0:
            piv[k] = 1
            r = int(ipiv[k]) - 1
            perm[k], perm[r] = perm[r], perm[k]
            k += 1
        else:
            piv[k], piv[k + 1] = 2, -2
            r = -int(ipiv[k]) - 1
            perm[k + 1], perm[r] = perm[r], perm[k + 1]
            k += 2
    return np.array(perm, dtype=np.int32), np.array(piv, dtype=np.int32)


def min_margin(trace) -> float:
    """Smallest relative margin over all pivot decisions (0 = a tie)."""
    alpha = 0.6403882032022076
    m = np.inf
    for d in trace:
        if d.kind == 0:
            continue
        if d.omega > 0:
            m = min(m, (d.omega - d.omega2) / d.omega)
            a, b = d.akk, alpha * d.omega
            m = min(m, abs(a - b) / max(a + b, 1e-300))
        if d.kind in (3, -1, 2):
            a, b = d.akk * d.sigma, alpha * d.omega * d.omega
            m = min(m, abs(a - b) / max(a + b, 1e-300))
            if d.kind != 3:
                a, b = d.arr, alpha * d.sigma
                m = min(m, abs(a - b) / max(a + b, 1e-300))
    return float(m)

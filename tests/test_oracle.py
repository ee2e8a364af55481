"""CPU tests that pin the oracle (oracle/bk_oracle.c) before it is trusted.

Pins, in order of strength:
* bitwise equality with the reference compiled from its own sources
  (oracle/_ref), when that checker has been built;
* the committed golden vectors written by the reference
  (tests/golden/reference_factors.npz, made by tests/golden/make_golden.py);
* the reference's own known-answer tests
  (tests/block_ldlt_decomposition_test.cc:352-508);
* mathematics: exact rational Bunch-Kaufman on tiny matrices, LAPACK
  dsytrf pivots on well-separated matrices, the unpivoted closed form on
  diagonally dominant matrices, pure 2x2 structure, reconstruction and
  solve residuals.
"""
import os
from fractions import Fraction

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_factors.npz")
KAT = np.array([[6.0, 12.0, 3.0, -6.0], [12.0, -8.0, -13.0, 4.0],
                [3.0, -13.0, -1.0, 1.0], [-6.0, 4.0, 1.0, 6.0]])
TOL = 1.0e-10  # tests/block_ldlt_decomposition_test.cc:31


def test_reference_compact_storage_kat(oracle):
    """TestReferenceCompactStorage, tests/block_ldlt_decomposition_test.cc:421-450."""
    f = oracle.dsytrf(KAT, 2)
    expected = np.array([[6.0, 12.0, -0.6875, 0.0], [0.0, -8.0, 0.59375, -0.5],
                         [0.0, 0.0, 8.78125, -0.6263345195729537],
                         [0.0, 0.0, 0.0, 4.555160142348754]])
    err = np.linalg.norm(np.triu(f.compact) - expected) / max(np.linalg.norm(expected), 1)
    assert err < TOL
    assert list(f.piv) == [2, -2, 1, 1]


def test_kernel_shapes_kat(oracle):
    """TestKernelShapes, tests/block_ldlt_decomposition_test.cc:367-385."""
    import ctypes
    L = oracle.lib()
    src = np.array([2.0, 4.0, 6.0]); dst = np.zeros(3)
    L.bko_kernel_scale_1x1(3, 2.0, oracle._d(src), 1, oracle._d(dst), 1)
    assert dst[0] == 1.0 and dst[2] == 3.0
    s1 = np.array([1.0, 2.0]); s2 = np.array([3.0, 4.0]); d1 = np.zeros(2); d2 = np.zeros(2)
    L.bko_kernel_scale_2x2(2, 2.0, 1.0, 3.0, oracle._d(s1), 1, oracle._d(s2), 1,
                           oracle._d(d1), 1, oracle._d(d2), 1)
    assert abs(d1[0]) < 1e-14 and abs(d2[0] - 1.0) < 1e-14


@pytest.mark.parametrize("case", [0, 1, 2])
@pytest.mark.parametrize("bs", [1, 2, 3, 4])
def test_fixed_factorizations(oracle, case, bs):
    """TestFixedFactorizations, tests/block_ldlt_decomposition_test.cc:476-508."""
    mats = [KAT,
            np.array([[0, 1, 0, 0], [1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], float),
            np.diag([2.0, -3.0, 0.0, 5.0])]
    A = mats[case]
    f = oracle.dsytrf(A, bs)
    assert oracle.relative_residual(A, f) < TOL


def test_random_factorizations(oracle):
    """TestRandomFactorizations, tests/block_ldlt_decomposition_test.cc:510-522
    (same seeds, same libstdc++ normal generator)."""
    worst = 0.0
    for n in range(3, 17):
        for bs in range(1, n + 1):
            for seed in range(4):
                A = oracle.gen_reference_symmetric(n, 10000 * n + 100 * bs + seed)
                f = oracle.dsytrf(A, bs)
                worst = max(worst, oracle.relative_residual(A, f))
    assert worst < TOL


def test_triangular_input_lower(oracle):
    """TestTriangularMatrixInput: the lower triangle describes the same
    matrix; the oracle factors the upper, so mirroring must agree."""
    for bs in range(1, 5):
        fu = oracle.dsytrf(KAT, bs)
        fl = oracle.dsytrf(np.tril(KAT).T.copy(), bs)
        assert np.array_equal(np.triu(fu.compact), np.triu(fl.compact))


def test_golden_vectors(oracle):
    """Bitwise equality with the reference's own outputs (committed)."""
    g = np.load(GOLDEN)
    meta = g["meta"]
    for idx, (n, bs, seed, scale) in enumerate(meta):
        n, bs = int(n), int(bs)
        iu = np.triu_indices(n)
        A = np.zeros((n, n))
        A[iu] = g["A_%d" % idx]
        A = A + np.triu(A, 1).T
        f = oracle.dsytrf(A, bs)
        assert np.array_equal(f.compact[iu], g["U_%d" % idx]), idx
        assert np.array_equal(f.perm, g["perm_%d" % idx]), idx
        assert np.array_equal(f.piv, g["piv_%d" % idx]), idx
        assert oracle.relative_residual(A, f) == g["res_%d" % idx][0], idx


def test_golden_generators(oracle):
    """The libstdc++-based generators reproduce the inputs the reference's
    own generators drew (fixtures written by RandomSymmetricMatrix,
    RandomSymmetricBlockMatrix and ResidualSamples compiled from the
    reference's test sources, tests/golden/make_golden.py)."""
    g = np.load(GOLDEN)
    meta = g["meta"]
    kinds = g["kinds"]
    checked = 0
    for idx, (n, bs, seed, scale) in enumerate(meta):
        if str(kinds[idx]).startswith("fixed"):
            continue
        n = int(n)
        A = oracle.gen_reference_symmetric(n, int(seed), float(scale))
        assert np.array_equal(A[np.triu_indices(n)], g["A_%d" % idx])
        checked += 1
    assert checked >= 30
    for t, (n, bs, seed, scale) in enumerate(g["block_meta"]):
        n = int(n)
        A = oracle.gen_stress_matrix(n, int(bs), int(seed), float(scale))
        assert np.array_equal(A[np.triu_indices(n)], g["block_A_%d" % t])
    expected = g["samples_5000_931300"]
    rows, cols = oracle.residual_samples(5000, 931300, expected.shape[1])
    assert expected.shape[1] == 4096
    assert np.array_equal(np.stack([rows, cols]), expected)


def test_golden_stress_case(oracle):
    """One whole large-stress case as the reference's own test code ran it
    (tests/block_ldlt_large_stress_test.cc:220-265): the oracle's compact
    factors of the same input are bitwise the reference's, and its sampled
    residual on them equals SampledRelativeResidual's value."""
    g = np.load(GOLDEN)
    n, bs, seed, scale = g["stress_meta"]
    n, bs, seed = int(n), int(bs), int(seed)
    A = oracle.gen_stress_matrix(n, bs, seed, float(scale))
    iu = np.triu_indices(n)
    for omp in (False, True):
        f = oracle.dsytrf(A, bs, omp=omp)
        assert np.array_equal(f.compact[iu], g["stress_U"])
        assert np.array_equal(f.perm, g["stress_perm"])
        assert np.array_equal(f.piv, g["stress_piv"])
        assert oracle.sampled_residual(A, f, seed) == g["stress_sampled_residual"][0]


def test_bitwise_against_reference_build(ref):
    """Bitwise equality with the reference library itself, including the
    workspace and the post-shift block geometry."""
    cases = []
    for n in (5, 10, 20, 50):
        for bs in (1, 2, 3, 5, 8, 16, 21, 32):
            if bs <= n:
                cases.append((n, bs))
    cases += [(200, 1), (200, 16), (200, 64), (129, 128)]
    scales = [1e-12, 1e-6, 1.0, 1e6, 1e12]
    for ci, (n, bs) in enumerate(cases):
        sc = scales[ci % len(scales)]
        A = ref.gen_reference_symmetric(n, 910000 + 10000 * n + 100 * bs + ci, sc)
        f = ref.dsytrf(A, bs)
        r = ref.ref_dsytrf(A, bs, full=True)
        assert np.array_equal(np.triu(f.compact), np.triu(r.compact)), (n, bs)
        assert np.array_equal(f.perm, r.perm) and np.array_equal(f.piv, r.piv), (n, bs)
        assert np.array_equal(np.triu(f.workspace), np.triu(r.workspace)), (n, bs)
        assert ref.relative_residual(A, f) == r.residual


@pytest.mark.parametrize("n,bs", [(400, 64), (500, 8), (500, 128), (1000, 128)])
def test_omp_oracle_bitwise_against_reference_build(ref, n, bs):
    """The OpenMP form of the oracle (used for n >= 400) against the
    reference's OpenMP build at the extended stress sizes and block sizes
    of tests/block_ldlt_decomposition_test.cc:45-58 (seeds as its
    BuildStressCases, :67-89, base 920000); both parallelize element-wise
    loops whose per-element summation order is fixed, so results are
    bitwise equal."""
    si = (n // 100 + bs) % 9
    scales = [1e-12, 1e-9, 1e-6, 1e-3, 1.0, 1e3, 1e6, 1e9, 1e12]
    seed = 920000 + 10000 * n + 100 * bs + 10 * si
    A = ref.ref_random_symmetric_matrix(n, seed, scales[si])
    _, r = ref.ref_time_factorize(A, bs, want_factors=True)
    f = ref.dsytrf(A, bs, omp=True)
    assert np.array_equal(np.triu(f.compact), np.triu(r.compact))
    assert np.array_equal(f.perm, r.perm) and np.array_equal(f.piv, r.piv)
    assert ref.relative_residual(A, f) < TOL


def test_generators_against_reference_build(ref):
    """The oracle's generators against the reference's own test helpers,
    compiled from its test sources: RandomSymmetricMatrix
    (tests/block_ldlt_decomposition_test.cc:148-161),
    RandomSymmetricBlockMatrix and ResidualSamples
    (tests/block_ldlt_large_stress_test.cc:116-136,204-218)."""
    for n, seed, scale in ((1, 3, 1.0), (17, 920007, 1e-6), (64, 5, 1e12), (131, 99, 1.0)):
        assert np.array_equal(ref.gen_reference_symmetric(n, seed, scale),
                              ref.ref_random_symmetric_matrix(n, seed, scale))
    for n, bs, seed, scale in ((1, 1, 0, 1.0), (50, 7, 931001, 1e3), (129, 128, 931002, 1.0),
                               (200, 64, 5, 1e-9)):
        assert np.array_equal(ref.gen_stress_matrix(n, bs, seed, scale),
                              ref.ref_random_symmetric_block_matrix(n, bs, seed, scale))
    for n, seed in ((1, 0), (2, 7), (5000, 931300), (10000, 940000), (65536, 12)):
        rows, cols = ref.ref_residual_samples(n, seed)
        r2, c2 = ref.residual_samples(n, seed, len(rows))
        assert np.array_equal(rows, r2) and np.array_equal(cols, c2)


@pytest.mark.parametrize("n,bs,scale", [(96, 16, 1.0), (160, 32, 1e6), (250, 64, 1e-6)])
def test_sampled_residual_against_reference_build(ref, n, bs, scale):
    """bko_sampled_residual equals the stress test's SampledRelativeResidual
    (tests/block_ldlt_large_stress_test.cc:160-244) on the same factors,
    bit for bit; also with a planted error in the factors, where the two
    must still agree (the reference's value comes from the same factors
    only when they are unmodified, so the planted case compares against
    the exact formula evaluated with numpy)."""
    seed = 931000 + n
    f_ref, sres = ref.ref_stress_case(n, bs, seed, scale)
    A = ref.gen_stress_matrix(n, bs, seed, scale)
    f = ref.dsytrf(A, bs)
    assert np.array_equal(np.triu(f.compact), np.triu(f_ref.compact))
    assert np.array_equal(f.perm, f_ref.perm) and np.array_equal(f.piv, f_ref.piv)
    assert ref.sampled_residual(A, f, seed) == sres
    # a planted error in an L entry of a sampled row moves both the same way
    rows, cols = ref.residual_samples(n, seed)
    inv = np.argsort(f.perm)
    i = int(inv[rows[10]])
    p = max(0, i - 1)
    if p < i and f.piv[p] != 2:
        bad = ref.Factorization(n, bs, f.compact.copy(), f.perm, f.piv)
        bad.compact[p, i] += 0.25 * max(1.0, abs(bad.compact[p, i]))
        L, D = ref.extract_ld(bad)
        R = L @ D @ L.T
        rec = R[inv[rows], inv[cols]]
        exp = A[rows, cols]
        want = np.linalg.norm(rec - exp) / max(np.linalg.norm(exp), 1.0)
        got = ref.sampled_residual(A, bad, seed)
        assert got > 1e-6 and abs(got - want) <= 1e-12 * max(1.0, want)


KERNEL_SHAPES = [(1, 1, 1), (7, 5, 3), (16, 16, 16), (33, 9, 65), (5, 40, 128), (0, 4, 4),
                 (4, 0, 4), (4, 4, 0)]


@pytest.mark.parametrize("rows,cols,inner", KERNEL_SHAPES)
def test_kernel_update_trailing_against_reference_build(ref, rows, cols, inner):
    """bko_kernel_update_trailing against kernels::UpdateTrailingBlock
    (src/kernels.cc:67-91, non-MKL) on strided operands, bit for bit;
    non-positive sizes leave the target untouched."""
    rng = np.random.default_rng(rows * 10000 + cols * 100 + inner)
    ls, ws, ts = max(rows, 1) + 3, max(cols, 1) + 5, max(cols, 1) + 2
    left = rng.standard_normal(max(inner, 1) * ls)
    work = rng.standard_normal(max(inner, 1) * ws)
    left[::7] = 0.0
    t0 = rng.standard_normal(max(rows, 1) * ts)
    a, b = t0.copy(), t0.copy()
    ref.lib().bko_kernel_update_trailing(rows, cols, inner, ref._d(left), ls, ref._d(work), ws,
                                         ref._d(a), ts)
    ref.ref().ref_kernel_update_trailing(rows, cols, inner, ref._d(left), ls, ref._d(work), ws,
                                         ref._d(b), ts)
    assert np.array_equal(a, b)
    if min(rows, cols, inner) <= 0:
        assert np.array_equal(a, t0)


@pytest.mark.parametrize("pivot", [2.5, -1e-300, 0.0, -0.0, 1e-17, np.inf])
@pytest.mark.parametrize("count,ss,ds", [(1, 1, 1), (9, 3, 2), (64, 1, 5), (0, 1, 1), (-2, 1, 1)])
def test_kernel_scale_1x1_against_reference_build(ref, pivot, count, ss, ds):
    """bko_kernel_scale_1x1 against kernels::ScaleOneByOnePivot
    (src/kernels.cc:11-31): division, the exact-zero pivot (zeros written),
    strides, non-positive counts (no-op); bit for bit."""
    rng = np.random.default_rng(count + 17)
    src = rng.standard_normal(max(count, 1) * ss)
    src[::4] = 0.0
    a = np.full(max(count, 1) * ds, 7.0)
    b = a.copy()
    ref.lib().bko_kernel_scale_1x1(count, pivot, ref._d(src), ss, ref._d(a), ds)
    ref.ref().ref_kernel_scale_1x1(count, pivot, ref._d(src), ss, ref._d(b), ds)
    assert np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("abc", [(2.0, 1.0, 3.0), (0.0, 1.0, 0.0), (1.0, 1.0, 1.0), (4.0, 2.0, 1.0),
                                 (1e-200, 1e-200, 1e-200), (-3.0, 0.5, 7.0)])
@pytest.mark.parametrize("count,s1,s2,d1,d2", [(1, 1, 1, 1, 1), (13, 2, 3, 1, 4), (0, 1, 1, 1, 1)])
def test_kernel_scale_2x2_against_reference_build(ref, abc, count, s1, s2, d1, d2):
    """bko_kernel_scale_2x2 against kernels::ScaleTwoByTwoPivot
    (src/kernels.cc:33-65): the 2x2 solve, an exactly singular block
    (det = 0: zeros written, including the underflowing det of 1e-200
    entries), strides and an empty count; bit for bit."""
    a_, b_, c_ = abc
    rng = np.random.default_rng(count * 3 + 1)
    x1 = rng.standard_normal(max(count, 1) * s1)
    x2 = rng.standard_normal(max(count, 1) * s2)
    A1, A2 = np.full(max(count, 1) * d1, -5.0), np.full(max(count, 1) * d2, -6.0)
    B1, B2 = A1.copy(), A2.copy()
    ref.lib().bko_kernel_scale_2x2(count, a_, b_, c_, ref._d(x1), s1, ref._d(x2), s2,
                                   ref._d(A1), d1, ref._d(A2), d2)
    ref.ref().ref_kernel_scale_2x2(count, a_, b_, c_, ref._d(x1), s1, ref._d(x2), s2,
                                   ref._d(B1), d1, ref._d(B2), d2)
    assert np.array_equal(A1, B1, equal_nan=True) and np.array_equal(A2, B2, equal_nan=True)


def test_reference_lower_input(ref):
    """BlockMatrix(..., MatrixTriangle::kLower) reads the lower triangle
    (src/block_matrix.cc:30-35); sentinel 12345 in the ignored triangle as in
    tests/block_ldlt_decomposition_test.cc:163-176."""
    A = KAT.copy()
    low = np.tril(A) + np.triu(np.full_like(A, 12345.0), 1)
    for bs in range(1, 5):
        r = ref.ref_dsytrf(low, bs, full=True, lower=True)
        assert r.residual < TOL
        f = ref.dsytrf(A, bs)
        assert np.array_equal(np.triu(f.compact), np.triu(r.compact))


# ---------------------------------------------------------------- mathematics

def exact_bk(A, alpha=0.6403882032022076):
    """Unblocked Bunch-Kaufman in exact rational arithmetic (LAPACK 'L'
    decisions); alpha enters only through comparisons of rationals with its
    binary64 value.  Returns (perm, piv) in the reference's encoding."""
    n = len(A)
    M = [[Fraction(x) for x in row] for row in A]
    a = Fraction(alpha)
    perm = list(range(n))
    piv = [0] * n
    k = 0

    def swap(i, j):
        M[i], M[j] = M[j], M[i]
        for row in M:
            row[i], row[j] = row[j], row[i]
        perm[i], perm[j] = perm[j], perm[i]

    while k < n:
        akk = abs(M[k][k])
        omega, r = Fraction(0), -1
        for j in range(k + 1, n):
            if abs(M[j][k]) > omega:
                omega, r = abs(M[j][k]), j
        if akk >= a * omega or omega == 0:
            size = 1
        else:
            sigma = max(abs(M[i][r]) for i in range(k, n) if i != r)
            if akk * sigma >= a * omega * omega:
                size = 1
            elif abs(M[r][r]) >= a * sigma:
                swap(k, r)
                size = 1
            else:
                swap(k + 1, r)
                size = 2
        if size == 1:
            piv[k] = 1
            d = M[k][k]
            if d != 0:
                for i in range(k + 1, n):
                    for j in range(k + 1, n):
                        M[i][j] -= M[i][k] * M[k][j] / d
        else:
            piv[k], piv[k + 1] = 2, -2
            D = [[M[k][k], M[k][k + 1]], [M[k + 1][k], M[k + 1][k + 1]]]
            det = D[0][0] * D[1][1] - D[0][1] * D[1][0]
            Di = [[D[1][1] / det, -D[0][1] / det], [-D[1][0] / det, D[0][0] / det]]
            for i in range(k + 2, n):
                li = [M[i][k] * Di[0][0] + M[i][k + 1] * Di[1][0],
                      M[i][k] * Di[0][1] + M[i][k + 1] * Di[1][1]]
                for j in range(k + 2, n):
                    M[i][j] -= li[0] * M[k][j] + li[1] * M[k + 1][j]
        k += size
    return perm, piv


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8])
def test_exact_rational_decisions(oracle, n):
    """Brute force on tiny matrices: the oracle's pivot decisions equal exact
    Bunch-Kaufman whenever every decision margin is clear of rounding."""
    rng = np.random.default_rng(100 + n)
    checked = 0
    for trial in range(25):
        A = rng.integers(-9, 10, size=(n, n)).astype(float)
        A = np.triu(A) + np.triu(A, 1).T
        for bs in sorted({1, 2, n}):
            f = oracle.dsytrf(A, bs, trace=True)
            if oracle.min_margin(f.trace) < 1e-9:
                continue
            perm, piv = exact_bk(A.tolist())
            assert list(f.piv) == piv, (A, bs)
            assert list(f.perm) == perm, (A, bs)
            checked += 1
    assert checked > 10


def test_lapack_dsytrf_pivots(oracle):
    """Same interchanges as LAPACK dsytrf('L') (scipy/OpenBLAS) on matrices
    with well-separated decisions, mapped as in
    tests/block_ldlt_decomposition_test.cc:196-240."""
    from scipy.linalg import lapack
    compared = 0
    for n in (16, 40, 96, 200):
        for seed in range(4):
            A = oracle.gen_synth_symmetric(n, 7000 + seed)
            f = oracle.dsytrf(A, 32, trace=True)
            if oracle.min_margin(f.trace) < 1e-8:
                continue
            _, ipiv, info = lapack.dsytrf(A, lower=1)
            perm, piv = oracle.lapack_to_reference(ipiv)
            assert np.array_equal(piv, f.piv), (n, seed)
            assert np.array_equal(perm, f.perm), (n, seed)
            compared += 1
    assert compared >= 8


def test_diagonally_dominant_closed_form(oracle):
    """No pivoting happens on a diagonally dominant matrix, so the factors are
    the textbook unpivoted LDL^T."""
    rng = np.random.default_rng(5)
    n = 30
    A = rng.uniform(-1, 1, (n, n))
    A = A + A.T
    A[np.diag_indices(n)] = np.where(rng.random(n) < 0.5, -1, 1) * (2 * n + rng.random(n))
    f = oracle.dsytrf(A, 7)
    assert np.all(f.piv == 1) and np.array_equal(f.perm, np.arange(n))
    L = np.eye(n)
    D = np.zeros(n)
    S = A.copy()
    for k in range(n):
        D[k] = S[k, k]
        L[k + 1:, k] = S[k + 1:, k] / D[k]
        S[k + 1:, k + 1:] -= np.outer(L[k + 1:, k], S[k, k + 1:])
    assert np.allclose(np.diag(f.compact), D, rtol=1e-13)
    assert np.allclose(np.triu(f.compact, 1), np.triu(L.T, 1), rtol=1e-12, atol=1e-14)


def test_pure_2x2_structure(oracle):
    """[[0,1],[1,0]] blocks force 2x2 pivots with D = the block itself."""
    n = 8
    A = np.zeros((n, n))
    for k in range(0, n, 2):
        A[k, k + 1] = A[k + 1, k] = 1.0 + k
    for bs in (1, 2, 3, 8):
        f = oracle.dsytrf(A, bs)
        assert list(f.piv) == [2, -2] * (n // 2)
        for k in range(0, n, 2):
            assert f.compact[k, k] == 0 and f.compact[k, k + 1] == 1.0 + k
        assert oracle.relative_residual(A, f) == 0.0


def test_sampled_residual_matches_full(oracle):
    """The sampled residual (tests/block_ldlt_large_stress_test.cc:220-244)
    equals the exact formula evaluated with numpy on the same samples."""
    n, seed = 60, 931000
    A = oracle.gen_stress_matrix(n, 8, seed, 1.0)
    f = oracle.dsytrf(A, 8)
    L, D = oracle.extract_ld(f)
    R = L @ D @ L.T
    inv = np.argsort(f.perm)
    rows, cols = oracle.residual_samples(n, seed)
    rec = R[inv[rows], inv[cols]]
    exp = A[rows, cols]
    want = np.linalg.norm(rec - exp) / max(np.linalg.norm(exp), 1.0)
    got = oracle.sampled_residual(A, f, seed)
    assert abs(got - want) <= 1e-13
    assert got < 1e-12


def test_solve(oracle):
    """Solve with the compact factors against numpy."""
    n = 150
    A = oracle.gen_synth_symmetric(n, 99)
    f = oracle.dsytrf(A, 16)
    b = np.random.default_rng(1).standard_normal((n, 3))
    x = oracle.dsytrs(f, b)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-10)
    r = b - A @ x
    eta = np.abs(r).max(0) / (np.abs(A).sum(1).max() * np.abs(x).max(0) + np.abs(b).max(0))
    assert eta.max() < 1e-14


def test_synth_generator_formula(oracle):
    """SplitMix64 keyed uniforms: exact binary64 values in [-1, 1)."""
    import struct
    def sm(x):
        x = (x + 0x9E3779B97F4A7C15) & (2**64 - 1)
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        return x ^ (x >> 31)
    for (seed, i, j) in [(1, 0, 0), (7, 3, 11), (2**40 + 3, 1000, 4095)]:
        key = sm(sm(seed) ^ ((i << 32) | j))
        want = 2.0 * ((key >> 11) * 2.0**-53) - 1.0
        assert oracle.synth_uniform(seed, i, j) == want

"""GPU tests of the C++ drop-in layer (namespace naive_block_ldlt).

* The reference's own test programs (tests/block_ldlt_decomposition_test.cc
  and tests/block_ldlt_large_stress_test.cc), compiled unchanged against
  include/naive_block_ldlt and linked with the GPU library by
  ``make -C oracle dropin``, must pass: every KAT, the 532 random and 387
  stress factorizations (< 1e-10), the n = 5000 sampled residuals (< 1e-8).
* tests/cpp/drop_in_dump drives the same API on the golden inputs; block
  geometry after 2x2 boundary shifts must equal the reference's exactly,
  pivots must agree on well-separated inputs, factors to 1e-9, residual
  within 10x of the reference's.
"""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_TESTS = os.path.join(ROOT, "oracle", "_ref", "ref_tests_on_dropin")
REF_STRESS = os.path.join(ROOT, "oracle", "_ref", "ref_stress_on_dropin")
DUMP = os.path.join(ROOT, "blockedbunchkaufmanldlt_amd", "_build", "drop_in_dump")
GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_factors.npz")


def run_dump(A, bs, lower=False, full=True):
    n = A.shape[0]
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("4i", n, bs, int(lower), int(full)))
            f.write(np.ascontiguousarray(A, dtype=np.float64).tobytes())
        subprocess.run([DUMP, fin, fout], check=True, timeout=120)
        raw = open(fout, "rb").read()
    off = 0
    status = struct.unpack_from("i", raw, off)[0]; off += 4
    assert status == 0
    U = np.frombuffer(raw, np.float64, n * n, off).reshape(n, n); off += 8 * n * n
    perm = np.frombuffer(raw, np.int32, n, off); off += 4 * n
    piv = np.frombuffer(raw, np.int32, n, off); off += 4 * n
    nb = (n + bs - 1) // bs
    nbk = nb * (nb + 1) // 2
    geom = np.frombuffer(raw, np.int32, 5 * nbk, off).reshape(nbk, 5); off += 20 * nbk
    res = struct.unpack_from("d", raw, off)[0]; off += 8
    W = np.frombuffer(raw, np.float64, n * n, off).reshape(n, n)
    return U, perm.astype(np.int32), piv.astype(np.int32), geom, res, W


@pytest.mark.skipif(not os.path.exists(REF_TESTS), reason="reference tests not built (make -C oracle dropin)")
def test_reference_unit_tests_on_dropin():
    r = subprocess.run([REF_TESTS], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "C++ block LDLT tests passed" in r.stdout


@pytest.mark.skipif(not os.path.exists(REF_STRESS), reason="reference tests not built (make -C oracle dropin)")
@pytest.mark.parametrize("mode", ["--self-test", "--default-large"])
def test_reference_stress_on_dropin(mode):
    r = subprocess.run([REF_STRESS, mode], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "C++ large stress tests passed" in r.stdout


def test_dropin_against_golden(oracle):
    g = np.load(GOLDEN)
    for idx, (n, bs, seed, scale) in enumerate(g["meta"]):
        n, bs = int(n), int(bs)
        iu = np.triu_indices(n)
        A = np.zeros((n, n))
        A[iu] = g["A_%d" % idx]
        A = A + np.triu(A, 1).T
        U, perm, piv, geom, res, W = run_dump(A, bs)
        o = oracle.dsytrf(A, bs, trace=True)
        assert res < 1e-10 and res <= 10 * max(g["res_%d" % idx][0], 1e-15), idx
        if oracle.min_margin(o.trace) > 1e-8:
            assert np.array_equal(piv, g["piv_%d" % idx]), idx
            assert np.array_equal(perm, g["perm_%d" % idx]), idx
            assert np.array_equal(geom, g["geom_%d" % idx]), idx
            Ug = g["U_%d" % idx]
            assert np.abs(U[iu] - Ug).max() <= 1e-9 * max(1.0, np.abs(Ug).max()), idx


def test_dropin_lower_triangle_and_compact_mode(oracle):
    A = oracle.gen_reference_symmetric(37, 4242)
    low = np.tril(A) + np.triu(np.full_like(A, 12345.0), 1)
    U1, p1, v1, g1, r1, _ = run_dump(A, 5)
    U2, p2, v2, g2, r2, _ = run_dump(low, 5, lower=True)
    assert r1 < 1e-10 and r2 < 1e-10
    assert np.array_equal(U1, U2) and np.array_equal(p1, p2) and np.array_equal(g1, g2)
    U3, p3, v3, g3, r3, _ = run_dump(A, 5, full=False)
    assert r3 == -1.0 and np.array_equal(U1, U3)


def test_dropin_shift_geometry_many(oracle):
    """Small block sizes force many 2x2 boundary shifts; the replayed block
    geometry must equal the one the oracle's panel boundaries imply (the
    oracle's geometry equals the reference's, tests/test_oracle.py)."""
    checked = 0
    for n in (9, 16, 31):
        for bs in (1, 2, 3, 4):
            A = oracle.gen_reference_symmetric(n, 10000 * n + 100 * bs + 7)
            o = oracle.dsytrf(A, bs, trace=True)
            if oracle.min_margin(o.trace) < 1e-8:
                continue
            U, perm, piv, geom, res, W = run_dump(A, bs)
            assert np.array_equal(piv, o.piv)
            rows = geom[[J * (J + 1) // 2 + J for J in range((n + bs - 1) // bs)]]
            assert np.array_equal(rows[:, 0], o.panel_start)
            assert np.array_equal(rows[:, 0] + rows[:, 2], o.panel_end)
            checked += 1
    assert checked > 5

"""CPU checks of the shipped libraries (no GPU, no compute calls).

* libbbk_ldlt.so loads and exports every function include/bbk_ldlt.h
  declares.
* libnaive_block_ldlt.so exports the C++ drop-in API of
  include/naive_block_ldlt/*.h (the names the reference's own test
  programs link against).
* The Python wrapper refuses to run without the built library (no CPU
  fallback exists).
"""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "blockedbunchkaufmanldlt_amd")
HEADER = os.path.join(ROOT, "include", "bbk_ldlt.h")
LIB = os.path.join(PKG, "libbbk_ldlt.so")
CXXLIB = os.path.join(PKG, "libnaive_block_ldlt.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(bbk_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_api():
    names = declared_functions()
    for must in ("bbk_create", "bbk_destroy", "bbk_dsytrf", "bbk_ssytrf", "bbk_dsytrf_host",
                 "bbk_dsytrs", "bbk_ssytrs", "bbk_dsyberr", "bbk_dsysv", "bbk_dsysv_mixed",
                 "bbk_dsytrf_batched", "bbk_ssytrf_batched", "bbk_kernel_scale_1x1",
                 "bbk_kernel_scale_2x2", "bbk_kernel_update_trailing", "bbk_dresidual",
                 "bbk_dsampled_residual", "bbk_dreconstruct", "bbk_dreconstruct_host",
                 "bbk_residual_samples"):
        assert must in names, must


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built (python -m blockedbunchkaufmanldlt_amd.build)")
def test_residual_sample_positions_match_reference_fixture():
    """bbk_dsampled_residual's sample positions (drawn on the host) are the
    reference stress test's: checked against the fixture written from the
    reference's own generator (tests/block_ldlt_large_stress_test.cc:212-227)."""
    import numpy as np
    g = np.load(os.path.join(ROOT, "tests", "golden", "reference_factors.npz"))
    expected = g["samples_5000_931300"]
    count = expected.shape[1]
    lib = ctypes.CDLL(LIB)
    rows = (ctypes.c_int * count)()
    cols = (ctypes.c_int * count)()
    assert lib.bbk_residual_samples(5000, 931300, count, rows, cols) == 0
    assert np.array_equal(np.stack([np.array(rows), np.array(cols)]), expected)
    assert lib.bbk_residual_samples(5000, 1, 3, rows, cols) == -3


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built (python -m blockedbunchkaufmanldlt_amd.build)")
def test_c_abi_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [name for name in declared_functions() if not hasattr(lib, name)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(CXXLIB), reason="drop-in library not built")
def test_dropin_exports_reference_api():
    out = subprocess.run(["nm", "-DC", "--defined-only", CXXLIB], capture_output=True,
                         text=True, check=True).stdout
    for sym in ("naive_block_ldlt::RegularMatrix::RegularMatrix(int, int)",
                "naive_block_ldlt::BlockMatrix::BlockMatrix(int, int)",
                "naive_block_ldlt::BlockMatrix::BlockMatrix(naive_block_ldlt::RegularMatrix const&, int, naive_block_ldlt::MatrixTriangle)",
                "naive_block_ldlt::BlockLdltDecomposition::Factorize()",
                "naive_block_ldlt::BlockLdltDecomposition::ReconstructOriginal() const",
                "naive_block_ldlt::kernels::ScaleOneByOnePivot(",
                "naive_block_ldlt::kernels::ScaleTwoByTwoPivot(",
                "naive_block_ldlt::kernels::UpdateTrailingBlock("):
        assert sym in out, sym


def test_wrapper_fails_loudly_without_library(monkeypatch):
    import blockedbunchkaufmanldlt_amd as B
    monkeypatch.setattr(B, "LIB_PATH", os.path.join(ROOT, "no_such_dir", "libbbk_ldlt.so"))
    monkeypatch.setattr(B, "_lib", None, raising=False)
    with pytest.raises(Exception):
        B.lib()

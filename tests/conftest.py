"""Shared test setup.

* ``gpu`` marks tests that need an MI355X; the CPU suite runs with
  ``-m "not gpu"``.
* The oracle (oracle/liboracle.so) is built on first use if missing; the
  reference checker (oracle/_ref/libbkref.so) is used when it has been
  built from the reference sources, otherwise those tests skip.
* GPU tests load the in-tree HIP library and fail loudly if it is absent.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")


def pytest_collection_modifyitems(config, items):
    """Run the tests that start GPU programs in subprocesses first.

    The panel and triangular-solve kernels are cooperative launches, which
    use the device's global wave sync resource; only one process at a time
    can hold it.  Once this pytest process has made a cooperative launch, a
    child process's cooperative launch is refused and the library falls
    back to an ordinary launch (tests/test_gpu_second_process.py covers
    that path), so the drop-in tests (tests/test_gpu_dropin.py, which run
    compiled programs) go before every test that launches in-process to
    exercise the cooperative path there too."""
    first = [it for it in items if it.module.__name__.endswith("test_gpu_dropin")]
    rest = [it for it in items if not it.module.__name__.endswith("test_gpu_dropin")]
    items[:] = first + rest


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ref(oracle):
    if not oracle.ref_available():
        pytest.skip("reference checker oracle/_ref not built (needs the reference tree)")
    oracle.ref()
    return oracle


@pytest.fixture(scope="session")
def bbk():
    import torch
    import blockedbunchkaufmanldlt_amd as B
    if not torch.cuda.is_available():
        raise RuntimeError("GPU test run without a visible GPU")
    B.lib()
    return B


@pytest.fixture(scope="session")
def handle(bbk):
    return bbk.Handle()

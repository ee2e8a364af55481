"""A second process on the same GPU.  Once a process has made a cooperative
launch, the runtime refuses cooperative launches from other processes on
the device.  The panel and triangular-sweep launches then fall back to an
ordinary launch: their grids never exceed the co-resident capacity, and
every wait is bounded, so a workgroup that was not scheduled would end the
call with BBK_ERR_HIP at the time limit, not hang the GPU.  Here this
process has already factored (a cooperative launch), and a child process
factors and solves on the same device and checks its own results."""
import os
import subprocess
import sys
import textwrap

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = textwrap.dedent("""
    import sys
    sys.path.insert(0, {root!r})
    import torch
    import blockedbunchkaufmanldlt_amd as B
    h = B.Handle()
    h.set_wait_limit(5000)  # a lost record ends the call within 5 s
    n = 1500
    A = h.gen_synth(n, 77, 1.0, 7, B.padded_ld(n))
    T = A.clone()
    perm, piv = h.sytrf(T)
    h.synchronize()
    rs = h.sampled_residual(A, T, perm, piv, 3)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    b = torch.randn((n, 2), generator=gen, dtype=torch.float64, device="cuda")
    X = torch.zeros_like(b)
    berr = h.sysv(A, b, X)
    h.synchronize()
    print("ok", rs, berr, flush=True)
""")


def test_second_process_factors_and_solves(handle):
    import torch
    import blockedbunchkaufmanldlt_amd as B
    # this process makes cooperative launches first
    n = 800
    A = handle.gen_synth(n, 78, 1.0, 7, B.padded_ld(n))
    T = A.clone()
    perm, piv = handle.sytrf(T)
    handle.synchronize()
    assert handle.sampled_residual(A, T, perm, piv, 1) < 1e-12
    print("parent factored; starting the child process", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD.format(root=ROOT)], capture_output=True,
                       text=True, timeout=110, cwd=ROOT)
    out = r.stdout.strip().splitlines()
    assert r.returncode == 0, r.stdout + r.stderr
    assert out and out[-1].startswith("ok"), r.stdout + r.stderr
    _, rs, berr = out[-1].split()
    assert float(rs) < 1e-12 and float(berr) < 1e-14, out[-1]
    del A, T
    torch.cuda.empty_cache()

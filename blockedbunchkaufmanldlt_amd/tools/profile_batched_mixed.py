"""Profiling driver: the benchmark's batched mixed-precision workload.

1024 matrices, n uniform in [256, 2048], A = Q diag(lambda) Q^T with a dense
random orthogonal Q and condition 10^U[4,7] per matrix, 4 right-hand sides,
tolerance 1e-12 (``workloads.c5_batch``), solved with
``Handle.sysv_mixed_batched`` ``reps`` times (default 2).  Each solve's wall
time and correction counts are printed, so a kernel trace of this script
(``rocprofv3 --kernel-trace --stats -- python3 <this file>``) can be read
beside them.  It makes no cooperative launches.
"""
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))

import torch  # noqa: E402

import blockedbunchkaufmanldlt_amd as B  # noqa: E402
from blockedbunchkaufmanldlt_amd import workloads  # noqa: E402


def main(reps: int = 2) -> None:
    h = B.Handle()
    ns, mats, Bs = workloads.c5_batch(torch, B.padded_ld, 1024)
    Xs = [torch.zeros((n, 4), dtype=torch.float64, device="cuda") for n in ns]
    for rep in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = h.sysv_mixed_batched(mats, Bs, Xs, tol=1e-12, max_iter=30)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0)
        print(f"solve {rep}: {ms:.1f} ms, corrections mean {info['iters'].mean():.2f} "
              f"max {info['iters'].max()}, converged {info['converged']}", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 2)

"""GPU tests of the symmetric-workspace batched kernel (team mode 1): one
wavefront per matrix, a full symmetric copy of the matrix in a workspace
slice, every interchange an update of the logical -> physical index map.

The kernel does the arithmetic of the global-memory team kernel with MFMA
update tiles, element for element, so its factors, perm and piv are
compared bit for bit with that kernel's and with the resident kernel's
(same row products, same MFMA tiles), and every matrix against the oracle
as in test_gpu_batched.py.  Panel widths from 1 (every 2x2 that reaches a
panel's last column is deferred) to 32, orders 1..256, binary64 and
binary32, a grid capped below the batch size (each workgroup factors
several matrices in turn, reusing its LDS and workspace slice), the HIP
graph, and the refusal of widths above 32.
"""
import numpy as np
import pytest

from test_gpu_batched import check, host_matrix, width_for

pytestmark = pytest.mark.gpu

SYM = 1
NS = [1, 2, 3, 7, 16, 31, 32, 33, 47, 64, 65, 100, 127, 128, 129, 150, 176, 177, 200, 230,
      255, 256]


def run(handle, mats, dtype, widths=None, teams=None, tiles=None):
    """Factor copies of mats with the given per-class tables; returns
    (factors, perms, pivs) as host arrays."""
    import torch
    saved = (handle.batched_widths(dtype), handle.batched_team_sizes(dtype),
             handle.batched_tiles(dtype))
    try:
        if widths is not None:
            handle.set_batched_widths(dtype, widths)
        if teams is not None:
            handle.set_batched_team_sizes(dtype, teams)
        if tiles is not None:
            handle.set_batched_tiles(dtype, tiles)
        work = [m.clone() for m in mats]
        perms = [torch.empty(m.shape[0], dtype=torch.int32, device="cuda") for m in mats]
        pivs = [torch.empty(m.shape[0], dtype=torch.int32, device="cuda") for m in mats]
        handle.sytrf_batched(work, perms, pivs)
        handle.synchronize()
        F = [np.triu(w[:, :w.shape[0]].double().cpu().numpy()) for w in work]
        return F, [p.cpu().numpy() for p in perms], [v.cpu().numpy() for v in pivs]
    finally:
        handle.set_batched_widths(dtype, saved[0])
        handle.set_batched_team_sizes(dtype, saved[1])
        handle.set_batched_tiles(dtype, saved[2])


def same(a, b):
    for x, y, what in ((a[0], b[0], "factors"), (a[1], b[1], "perm"), (a[2], b[2], "piv")):
        for m, (u, v) in enumerate(zip(x, y)):
            assert np.array_equal(u, v), (what, m)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("bw,ts", [(1, 16), (3, 32), (8, 16), (16, 32), (24, 16), (32, 32)])
def test_sym_bitwise_equals_team_kernels(oracle, handle, precision, bw, ts):
    import torch
    dtype = torch.float64 if precision == "fp64" else torch.float32
    mats = handle.gen_synth_batched(NS, seed=6100 + 10 * bw + ts, scale=1.0, zero_diag_stride=7,
                                    dtype=dtype, pad=3)
    w = [bw] * 8
    sym = run(handle, mats, dtype, w, [SYM] * 4 + [256] * 4, [ts] * 8)
    team = run(handle, mats, dtype, w, [64] * 8, [16] * 8)
    same(sym, team)
    resident = run(handle, mats, dtype, w, [0] * 4 + [256] * 4, [ts] * 8)
    same(sym, resident)
    for m, T in enumerate(mats):
        A = host_matrix(T)
        g = oracle.Factorization(A.shape[0], bw, sym[0][m], sym[1][m].astype(np.int32),
                                 sym[2][m].astype(np.int32))
        check(oracle, A, g, bw, fp32=(precision == "fp32"))
    assert any((v == 2).any() for v in sym[2])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sym_several_matrices_per_workgroup(oracle, handle, precision):
    """Grid capped at 3 workgroups for 60 matrices: bitwise the uncapped
    result (each workgroup reuses its LDS, index map and workspace slice)."""
    import torch
    dtype = torch.float64 if precision == "fp64" else torch.float32
    rng = np.random.default_rng(17)
    ns = [int(v) for v in rng.integers(1, 257, size=60)] + [256]
    mats = handle.gen_synth_batched(ns, seed=333, scale=1.0, zero_diag_stride=5, dtype=dtype)
    teams = [SYM] * 4 + [256] * 4
    full = run(handle, mats, dtype, None, teams, [32] * 8)
    handle.set_batched_grid_limit(3)
    try:
        capped = run(handle, mats, dtype, None, teams, [32] * 8)
    finally:
        handle.set_batched_grid_limit(0)
    same(full, capped)
    widths = handle.batched_widths(dtype)
    for m, T in enumerate(mats[:20]):
        A = host_matrix(T)
        n = A.shape[0]
        g = oracle.Factorization(n, 0, full[0][m], full[1][m].astype(np.int32),
                                 full[2][m].astype(np.int32))
        check(oracle, A, g, width_for(widths, n), fp32=(precision == "fp32"))


def test_sym_graph_replay(handle):
    """A graph captured in this mode owns its workspace slices; replays with
    the inputs restored eagerly equal the eager run bit for bit."""
    import torch
    dtype = torch.float64
    ns = [40, 90, 130, 200, 256, 64, 1, 17]
    src = handle.gen_synth_batched(ns, seed=4040, zero_diag_stride=7, dtype=dtype)
    teams = [SYM] * 4 + [256] * 4
    eager = run(handle, src, dtype, None, teams, [16] * 8)
    saved = (handle.batched_team_sizes(dtype), handle.batched_tiles(dtype))
    handle.set_batched_team_sizes(dtype, teams)
    handle.set_batched_tiles(dtype, [16] * 8)
    try:
        work = [s.clone() for s in src]
        perms = [torch.empty(n, dtype=torch.int32, device="cuda") for n in ns]
        pivs = [torch.empty(n, dtype=torch.int32, device="cuda") for n in ns]
        g = handle.graph_sytrf_batched(work, perms, pivs)
        try:
            for rep in range(3):
                for w, s in zip(work, src):
                    w.copy_(s)
                torch.cuda.synchronize()
                g.launch()
                handle.synchronize()
                got = ([np.triu(w[:, :w.shape[0]].cpu().numpy()) for w in work],
                       [p.cpu().numpy() for p in perms], [v.cpu().numpy() for v in pivs])
                same(eager, got)
        finally:
            g.close()
    finally:
        handle.set_batched_team_sizes(dtype, saved[0])
        handle.set_batched_tiles(dtype, saved[1])


def test_sym_refuses_wide_panels(handle):
    import torch
    from blockedbunchkaufmanldlt_amd import BbkError
    mats = handle.gen_synth_batched([100, 50], seed=1, dtype=torch.float64)
    with pytest.raises(BbkError):
        run(handle, mats, torch.float64, [48] * 8, [SYM] * 4 + [256] * 4, [16] * 8)

"""MI355X (gfx950) blocked Bunch-Kaufman LDL^T: Python bindings.

Thin ctypes bindings over the C ABI in ``include/bbk_ldlt.h``
(``libbbk_ldlt.so``, built in-tree by ``blockedbunchkaufmanldlt_amd.build``).
Device buffers are torch tensors on the current HIP device; calls are
queued on torch's current stream.  There is no CPU fallback: every entry
point raises if the HIP library is missing or no GPU is present.

Layout: an order-n matrix is a row-major ``(n, lda)`` tensor whose upper
triangle (i <= j) holds A -- the byte layout of LAPACK's column-major
lower triangle.  Results use the reference's compact encoding (see the C
header): D on the diagonal, D(i,i+1) for 2x2 pivots, L(j,i) at (i,j)
otherwise, plus ``perm`` and ``piv`` int32 vectors.
"""
from __future__ import annotations

import contextlib
import ctypes
import os
from typing import Optional, Sequence, Tuple

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libbbk_ldlt.so")
CXXLIB_PATH = os.path.join(PKG, "libnaive_block_ldlt.so")

BBK_OK = 0
BBK_ERR_HIP = 1
BBK_ERR_ALLOC = 2
BBK_ERR_UNSUPPORTED = 3
BBK_ERR_NOT_CONVERGED = 4

_lib = None


class BbkError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bbk error {code}: {msg}")
        self.code = code


def lib() -> ctypes.CDLL:
    """Load libbbk_ldlt.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: run `python -m blockedbunchkaufmanldlt_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    ip = ctypes.POINTER(ctypes.c_int)
    L.bbk_create.argtypes = [ctypes.POINTER(vp), vp]
    L.bbk_destroy.argtypes = [vp]
    L.bbk_get_stream.argtypes = [vp, ctypes.POINTER(vp)]
    L.bbk_synchronize.argtypes = [vp]
    L.bbk_last_error.argtypes = [vp]
    L.bbk_last_error.restype = ctypes.c_char_p
    L.bbk_set_wait_limit.argtypes = [vp, i32]
    L.bbk_set_panel_threads.argtypes = [vp, i32]
    L.bbk_get_info.argtypes = [vp, ip, ip]
    for name in ("bbk_dsytrf", "bbk_ssytrf"):
        getattr(L, name).argtypes = [vp, i32, i32, vp, i64, vp, vp]
    L.bbk_dsytrf_host.argtypes = [vp, i32, i32, vp, i64, vp, vp]
    L.bbk_dsytrf_ooc.argtypes = [vp, i32, i32, vp, i64, vp, vp, ctypes.c_size_t]
    L.bbk_host_alloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    L.bbk_host_free.argtypes = [vp]
    for name in ("bbk_dsytrs", "bbk_ssytrs"):
        getattr(L, name).argtypes = [vp, i32, i32, vp, i64, vp, vp, vp, i64]
    L.bbk_dsyberr.argtypes = [vp, i32, i32, vp, i64, vp, i64, vp, i64, vp]
    L.bbk_dresidual.argtypes = [vp, i32, vp, i64, vp, i64, vp, vp, ctypes.POINTER(f64)]
    L.bbk_dreconstruct.argtypes = [vp, i32, vp, i64, vp, vp, vp, i64]
    L.bbk_dsytrf_to_lapack.argtypes = [vp, i32, vp, i64, vp, vp, vp]
    L.bbk_dsytrf_from_lapack.argtypes = [vp, i32, vp, i64, vp, vp, vp]
    L.bbk_dsampled_residual.argtypes = [vp, i32, vp, i64, vp, i64, vp, vp, i32, i32,
                                        ctypes.POINTER(f64)]
    L.bbk_dsysv_mixed.argtypes = [vp, i32, i32, vp, i64, vp, i64, vp, i64, f64,
                                  i32, ip, ctypes.POINTER(f64)]
    L.bbk_dsysv.argtypes = [vp, i32, i32, vp, i64, vp, i64, vp, i64,
                            ctypes.POINTER(f64)]
    for name in ("bbk_dsytrf_batched", "bbk_ssytrf_batched"):
        getattr(L, name).argtypes = [vp, i32, ip, vp, vp, vp, vp]
    L.bbk_batched_prepare.argtypes = [vp, i32, ip, i32]
    L.bbk_batched_get_widths.argtypes = [vp, i32, ip]
    L.bbk_batched_set_widths.argtypes = [vp, i32, ip]
    L.bbk_batched_tune.argtypes = [vp, i32, i32, ip]
    L.bbk_batched_set_grid_limit.argtypes = [vp, i32]
    L.bbk_batched_get_tiles.argtypes = [vp, i32, ip]
    L.bbk_batched_set_tiles.argtypes = [vp, i32, ip]
    L.bbk_dsysv_mixed_batched.argtypes = [vp, i32, ip, i32, vp, vp, vp, vp, vp, vp, f64, i32,
                                          vp, vp]
    L.bbk_batched_get_team_sizes.argtypes = [vp, i32, ip]
    L.bbk_batched_set_team_sizes.argtypes = [vp, i32, ip]
    for name in ("bbk_graph_create_dsytrf_batched", "bbk_graph_create_ssytrf_batched"):
        getattr(L, name).argtypes = [vp, i32, ip, vp, vp, vp, vp, ctypes.POINTER(vp)]
    L.bbk_graph_launch.argtypes = [vp]
    L.bbk_graph_destroy.argtypes = [vp]
    for name in ("bbk_dgen_synth_batched", "bbk_sgen_synth_batched"):
        getattr(L, name).argtypes = [vp, i32, vp, ctypes.c_uint64, f64, i32, vp, vp]
    L.bbk_kernel_scale_1x1.argtypes = [vp, i32, f64, vp, i32, vp, i32]
    L.bbk_kernel_scale_2x2.argtypes = [vp, i32, f64, f64, f64, vp, i32, vp, i32,
                                       vp, i32, vp, i32]
    L.bbk_kernel_update_trailing.argtypes = [vp, i32, i32, i32, vp, i32, vp, i32,
                                             vp, i32]
    L.bbk_dgen_synth.argtypes = [vp, i32, ctypes.c_uint64, f64, i32, vp, i64]
    L.bbk_set_timing.argtypes = [vp, i32]
    L.bbk_last_timing.argtypes = [vp, ctypes.POINTER(f64), ctypes.POINTER(f64),
                                  ctypes.POINTER(f64), ip]
    _lib = L
    return L


def _ptr(t) -> Optional[int]:
    return None if t is None else t.data_ptr()


class Handle:
    """A bbk handle ordered with torch's current stream on the current device.

    With a non-default torch stream the handle queues its work on that
    stream.  On torch's default stream the handle owns a non-blocking
    stream (cooperative launches are refused on streams tied to the null
    stream) and every call is fenced both ways with stream waits, so torch
    work before a call is visible to it and torch work after it sees its
    results."""

    def __init__(self, stream=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("blockedbunchkaufmanldlt_amd needs a HIP GPU")
        self._torch = torch
        st = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        rc = lib().bbk_create(ctypes.byref(h), ctypes.c_void_p(st.cuda_stream or None))
        if rc != 0:
            raise BbkError(rc, "bbk_create failed")
        self.h = h
        self._own = None
        if not st.cuda_stream:
            sp = ctypes.c_void_p()
            lib().bbk_get_stream(h, ctypes.byref(sp))
            self._own = torch.cuda.ExternalStream(sp.value)
        self.stream = self._own if self._own is not None else st

    @contextlib.contextmanager
    def ordered(self):
        """Fence a block of library calls against torch's current stream."""
        if self._own is None:
            yield
            return
        cur = self._torch.cuda.current_stream()
        self._own.wait_stream(cur)
        try:
            yield
        finally:
            cur.wait_stream(self._own)

    def close(self):
        if getattr(self, "h", None):
            lib().bbk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int, what: str):
        if rc != 0:
            msg = lib().bbk_last_error(self.h).decode()
            raise BbkError(rc, f"{what}: {msg}")

    def synchronize(self):
        self.check(lib().bbk_synchronize(self.h), "synchronize")

    def set_wait_limit(self, ms: int):
        """Time limit of a cooperative launch's waits in ms (0 = default,
        < 0 = every wait fails at once)."""
        self.check(lib().bbk_set_wait_limit(self.h, ms), "set_wait_limit")

    def set_panel_threads(self, threads: int):
        """Panel kernel variant: 0 = by order (256-column workgroups up to
        n = 65536, 512-column ones up to 131072), 256 or 512 = that one."""
        self.check(lib().bbk_set_panel_threads(self.h, threads), "set_panel_threads")

    def info(self):
        """(info, nonfinite) of the last factorization: info = k + 1 for the
        first singular pivot column k (0 = none), nonfinite = 1 when a NaN or
        infinity appeared in a pivot row."""
        i, nf = ctypes.c_int(), ctypes.c_int()
        self.check(lib().bbk_get_info(self.h, ctypes.byref(i), ctypes.byref(nf)), "get_info")
        return i.value, nf.value

    def set_timing(self, on: bool):
        lib().bbk_set_timing(self.h, int(on))

    def last_timing(self):
        p, u, s = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int()
        lib().bbk_last_timing(self.h, ctypes.byref(p), ctypes.byref(u),
                              ctypes.byref(s), ctypes.byref(n))
        return {"panel_ms": p.value, "update_ms": u.value, "swap_ms": s.value,
                "panels": n.value}

    # ---- factorization ------------------------------------------------
    def sytrf(self, A, nb: int = 0, perm=None, piv=None):
        """Factor A in place (float64 -> bbk_dsytrf, float32 -> bbk_ssytrf).
        Returns (perm, piv) int32 device tensors."""
        torch = self._torch
        n, lda = A.shape[0], A.stride(0)
        assert A.is_cuda and A.stride(1) == 1 and A.shape[1] >= n
        if perm is None:
            perm = torch.empty(n, dtype=torch.int32, device=A.device)
        if piv is None:
            piv = torch.empty(n, dtype=torch.int32, device=A.device)
        if A.dtype == torch.float64:
            with self.ordered():
                rc = lib().bbk_dsytrf(self.h, n, nb, _ptr(A), lda, _ptr(perm), _ptr(piv))
        elif A.dtype == torch.float32:
            with self.ordered():
                rc = lib().bbk_ssytrf(self.h, n, nb, _ptr(A), lda, _ptr(perm), _ptr(piv))
        else:
            raise TypeError("A must be float64 or float32")
        self.check(rc, "sytrf")
        return perm, piv

    def sytrf_ooc(self, M: "PinnedMatrix", nb: int = 0, device_bytes: int = 0):
        """Factor the float64 matrix held in pinned host memory M in place
        (bbk_dsytrf_ooc): HBM use for the matrix's working copies bounded by
        device_bytes (0 = the free HBM less 1 GiB).  Synchronous.  Returns
        (perm, piv) as int32 numpy arrays."""
        import numpy as np
        perm = np.empty(M.n, dtype=np.int32)
        piv = np.empty(M.n, dtype=np.int32)
        with self.ordered():
            rc = lib().bbk_dsytrf_ooc(self.h, M.n, nb, M.ptr, M.lda,
                                      perm.ctypes.data, piv.ctypes.data, device_bytes)
        self.check(rc, "sytrf_ooc")
        return perm, piv

    def sytrs(self, F, perm, piv, B):
        """Solve in place: B (n, nrhs) row-major -> X."""
        torch = self._torch
        n, lda = F.shape[0], F.stride(0)
        nrhs = B.shape[1] if B.dim() == 2 else 1
        ldb = B.stride(0) if B.dim() == 2 else 1
        fn = lib().bbk_dsytrs if F.dtype == torch.float64 else lib().bbk_ssytrs
        with self.ordered():
            self.check(fn(self.h, n, nrhs, _ptr(F), lda, _ptr(perm), _ptr(piv), _ptr(B), ldb),
                       "sytrs")
        return B

    def syberr(self, A, B, X):
        torch = self._torch
        n, lda = A.shape[0], A.stride(0)
        nrhs = B.shape[1] if B.dim() == 2 else 1
        ldb = B.stride(0) if B.dim() == 2 else 1
        ldx = X.stride(0) if X.dim() == 2 else 1
        eta = torch.empty(nrhs, dtype=torch.float64, device=A.device)
        with self.ordered():
            self.check(lib().bbk_dsyberr(self.h, n, nrhs, _ptr(A), lda, _ptr(B), ldb,
                                         _ptr(X), ldx, _ptr(eta)), "syberr")
        return eta

    def reconstruct(self, F, perm, piv):
        """P L D L^T P^T (n x n float64 on F's device) from compact factors."""
        n = F.shape[0]
        out = self._torch.empty((n, n), dtype=self._torch.float64, device=F.device)
        with self.ordered():
            self.check(lib().bbk_dreconstruct(self.h, n, _ptr(F), F.stride(0), _ptr(perm),
                                              _ptr(piv), _ptr(out), out.stride(0)),
                       "reconstruct")
        return out

    def to_lapack(self, F, perm, piv):
        """Convert compact factors in place to LAPACK dsytrf('L') form;
        returns ipiv (1-based, int32 on F's device)."""
        n = F.shape[0]
        ipiv = self._torch.empty(n, dtype=self._torch.int32, device=F.device)
        with self.ordered():
            self.check(lib().bbk_dsytrf_to_lapack(self.h, n, _ptr(F), F.stride(0), _ptr(perm),
                                                  _ptr(piv), _ptr(ipiv)), "to_lapack")
        return ipiv

    def from_lapack(self, F, ipiv):
        """Convert a LAPACK dsytrf('L') factor in place to this library's
        form; returns (perm, piv)."""
        n = F.shape[0]
        perm = self._torch.empty(n, dtype=self._torch.int32, device=F.device)
        piv = self._torch.empty(n, dtype=self._torch.int32, device=F.device)
        with self.ordered():
            self.check(lib().bbk_dsytrf_from_lapack(self.h, n, _ptr(F), F.stride(0), _ptr(ipiv),
                                                    _ptr(perm), _ptr(piv)), "from_lapack")
        return perm, piv

    def residual(self, A, F, perm, piv) -> float:
        """||P L D L^T P^T - A||_F / max(||A||_F, 1) on the device (A: the
        factored matrix's upper triangle; F, perm, piv from sytrf)."""
        out = ctypes.c_double()
        with self.ordered():
            self.check(lib().bbk_dresidual(self.h, A.shape[0], _ptr(A), A.stride(0), _ptr(F),
                                           F.stride(0), _ptr(perm), _ptr(piv),
                                           ctypes.byref(out)), "residual")
        return out.value

    def sampled_residual(self, A, F, perm, piv, seed: int, count: int = 4096) -> float:
        """The stress test's sampled relative residual (same sample positions
        for the same seed) on the device."""
        out = ctypes.c_double()
        with self.ordered():
            self.check(lib().bbk_dsampled_residual(self.h, A.shape[0], _ptr(A), A.stride(0),
                                                   _ptr(F), F.stride(0), _ptr(perm), _ptr(piv),
                                                   seed, count, ctypes.byref(out)),
                       "sampled residual")
        return out.value

    def sysv_mixed(self, A, B, X, tol: float = 1e-12, max_iter: int = 30):
        n, lda = A.shape[0], A.stride(0)
        nrhs = B.shape[1] if B.dim() == 2 else 1
        ldb = B.stride(0) if B.dim() == 2 else 1
        ldx = X.stride(0) if X.dim() == 2 else 1
        it = ctypes.c_int()
        be = ctypes.c_double()
        with self.ordered():
            rc = lib().bbk_dsysv_mixed(self.h, n, nrhs, _ptr(A), lda, _ptr(B), ldb, _ptr(X),
                                       ldx, tol, max_iter, ctypes.byref(it), ctypes.byref(be))
        if rc not in (0, BBK_ERR_NOT_CONVERGED):
            self.check(rc, "sysv_mixed")
        return {"iters": it.value, "berr": be.value, "converged": rc == 0}

    def sysv(self, A, B, X):
        n, lda = A.shape[0], A.stride(0)
        nrhs = B.shape[1] if B.dim() == 2 else 1
        ldb = B.stride(0) if B.dim() == 2 else 1
        ldx = X.stride(0) if X.dim() == 2 else 1
        be = ctypes.c_double()
        with self.ordered():
            self.check(lib().bbk_dsysv(self.h, n, nrhs, _ptr(A), lda, _ptr(B), ldb, _ptr(X),
                                       ldx, ctypes.byref(be)), "sysv")
        return be.value

    # ---- batched factorization ------------------------------------------
    def _batch_arrays(self, mats, perms, pivs):
        torch = self._torch
        dev = mats[0].device
        count = len(mats)
        ns = (ctypes.c_int * count)(*[m.shape[0] for m in mats])
        aptr = torch.tensor([m.data_ptr() for m in mats], dtype=torch.int64, device=dev)
        lds = torch.tensor([m.stride(0) for m in mats], dtype=torch.int64, device=dev)
        pp = torch.tensor([p.data_ptr() for p in perms], dtype=torch.int64, device=dev)
        vp = torch.tensor([p.data_ptr() for p in pivs], dtype=torch.int64, device=dev)
        return count, ns, (aptr, lds, pp, vp)

    def sytrf_batched(self, mats: Sequence, perms: Sequence, pivs: Sequence):
        """Factor a list of (n_i, lda_i) tensors (all float64 or all float32)
        in place in one persistent launch; perms/pivs: int32 tensors of n_i."""
        torch = self._torch
        count, ns, arrs = self._batch_arrays(mats, perms, pivs)
        fn = lib().bbk_dsytrf_batched if mats[0].dtype == torch.float64 else \
            lib().bbk_ssytrf_batched
        with self.ordered():
            self.check(fn(self.h, count, ns, *[_ptr(a) for a in arrs]), "sytrf_batched")
        # keep the pointer arrays alive until the queued work has run
        self._keep = arrs

    def graph_sytrf_batched(self, mats: Sequence, perms: Sequence, pivs: Sequence):
        """Capture the batched factorization of these buffers into a HIP graph;
        returns a BatchGraph whose launch() replays it on this handle's stream."""
        torch = self._torch
        count, ns, arrs = self._batch_arrays(mats, perms, pivs)
        torch.cuda.synchronize()
        g = ctypes.c_void_p()
        fn = lib().bbk_graph_create_dsytrf_batched if mats[0].dtype == torch.float64 else \
            lib().bbk_graph_create_ssytrf_batched
        self.check(fn(self.h, count, ns, *[_ptr(a) for a in arrs], ctypes.byref(g)),
                   "graph_sytrf_batched")
        return BatchGraph(g, arrs, self)

    def sysv_mixed_batched(self, mats: Sequence, Bs: Sequence, Xs: Sequence,
                           tol: float = 1e-12, max_iter: int = 30):
        """Batched FP32 factorization + FP64 refinement of A_m X_m = B_m
        (float64 (n, lda) matrices, upper triangle read, not modified;
        B_m, X_m float64 (n, nrhs)).  Returns a dict with per-matrix
        iterations and backward errors (numpy arrays) and `converged`."""
        torch = self._torch
        count = len(mats)
        dev = mats[0].device
        nrhs = Bs[0].shape[1]
        ns = (ctypes.c_int * count)(*[m.shape[0] for m in mats])
        def arr(vals):
            return torch.tensor(vals, dtype=torch.int64, device=dev)
        keep = [arr([m.data_ptr() for m in mats]), arr([m.stride(0) for m in mats]),
                arr([b.data_ptr() for b in Bs]), arr([b.stride(0) for b in Bs]),
                arr([x.data_ptr() for x in Xs]), arr([x.stride(0) for x in Xs])]
        iters = torch.zeros(count, dtype=torch.int32, device=dev)
        berr = torch.zeros(count, dtype=torch.float64, device=dev)
        with self.ordered():
            rc = lib().bbk_dsysv_mixed_batched(self.h, count, ns, nrhs,
                                               *[_ptr(k) for k in keep], tol, max_iter,
                                               _ptr(iters), _ptr(berr))
        if rc not in (0, BBK_ERR_NOT_CONVERGED):
            self.check(rc, "sysv_mixed_batched")
        return {"iters": iters.cpu().numpy(), "berr": berr.cpu().numpy(),
                "converged": rc == 0}

    def batched_prepare(self, ns: Sequence[int], dtype):
        eb = 8 if dtype == self._torch.float64 else 4
        arr = (ctypes.c_int * len(ns))(*ns)
        self.check(lib().bbk_batched_prepare(self.h, len(ns), arr, eb), "batched_prepare")

    def batched_widths(self, dtype):
        eb = 8 if dtype == self._torch.float64 else 4
        arr = (ctypes.c_int * 8)()
        self.check(lib().bbk_batched_get_widths(self.h, eb, arr), "batched_get_widths")
        return list(arr)

    def set_batched_widths(self, dtype, widths: Sequence[int]):
        eb = 8 if dtype == self._torch.float64 else 4
        arr = (ctypes.c_int * 8)(*widths)
        self.check(lib().bbk_batched_set_widths(self.h, eb, arr), "batched_set_widths")

    def batched_team_sizes(self, dtype):
        eb = 8 if dtype == self._torch.float64 else 4
        arr = (ctypes.c_int * 8)()
        self.check(lib().bbk_batched_get_team_sizes(self.h, eb, arr), "batched_get_team_sizes")
        return list(arr)

    def set_batched_team_sizes(self, dtype, sizes: Sequence[int]):
        eb = 8 if dtype == self._torch.float64 else 4
        arr = (ctypes.c_int * 8)(*sizes)
        self.check(lib().bbk_batched_set_team_sizes(self.h, eb, arr), "batched_set_team_sizes")

    def batched_tiles(self, dtype):
        eb = 8 if dtype == self._torch.float64 else 4
        arr = (ctypes.c_int * 8)()
        self.check(lib().bbk_batched_get_tiles(self.h, eb, arr), "batched_get_tiles")
        return list(arr)

    def set_batched_tiles(self, dtype, tiles: Sequence[int]):
        eb = 8 if dtype == self._torch.float64 else 4
        arr = (ctypes.c_int * 8)(*tiles)
        self.check(lib().bbk_batched_set_tiles(self.h, eb, arr), "batched_set_tiles")

    def set_batched_grid_limit(self, max_workgroups: int):
        """Cap the persistent batched kernels' workgroups (0 = no cap)."""
        self.check(lib().bbk_batched_set_grid_limit(self.h, max_workgroups),
                   "batched_set_grid_limit")

    def batched_tune(self, dtype, max_n: int):
        eb = 8 if dtype == self._torch.float64 else 4
        arr = (ctypes.c_int * 8)()
        self.check(lib().bbk_batched_tune(self.h, eb, max_n, arr), "batched_tune")
        return list(arr)

    def gen_synth_batched(self, ns: Sequence[int], seed: int, scale: float = 1.0,
                          zero_diag_stride: int = 0, dtype=None, pad: int = 0):
        """Matrix m = gen_synth(ns[m], seed + m, ...) (rounded to float32 for
        dtype float32); returns a list of (n, n + pad) tensors."""
        torch = self._torch
        dtype = dtype or torch.float64
        # zero-filled, so the padding columns and the unused lower triangle
        # hold defined values (the generator writes the n x n block only)
        mats = [torch.zeros((n, n + pad), dtype=dtype, device="cuda") for n in ns]
        dn = torch.tensor(list(ns), dtype=torch.int32, device="cuda")
        aptr = torch.tensor([m.data_ptr() for m in mats], dtype=torch.int64, device="cuda")
        lds = torch.tensor([m.stride(0) for m in mats], dtype=torch.int64, device="cuda")
        fn = lib().bbk_dgen_synth_batched if dtype == torch.float64 else \
            lib().bbk_sgen_synth_batched
        with self.ordered():
            self.check(fn(self.h, len(ns), _ptr(dn), seed, scale, zero_diag_stride, _ptr(aptr),
                          _ptr(lds)), "gen_synth_batched")
        self.synchronize()
        return mats

    def gen_synth(self, n: int, seed: int, scale: float = 1.0, zero_diag_stride: int = 0,
                  lda: Optional[int] = None, dtype=None):
        torch = self._torch
        lda = lda or n
        A = torch.empty((n, lda), dtype=torch.float64, device="cuda")
        with self.ordered():
            self.check(lib().bbk_dgen_synth(self.h, n, seed, scale, zero_diag_stride,
                                            _ptr(A), lda), "gen_synth")
        return A

    # ---- kernels.h counterparts -----------------------------------------
    def kernel_scale_1x1(self, count, pivot, src, src_stride, dst, dst_stride):
        with self.ordered():
            self.check(lib().bbk_kernel_scale_1x1(self.h, count, pivot, _ptr(src), src_stride,
                                                  _ptr(dst), dst_stride), "scale_1x1")

    def kernel_scale_2x2(self, count, a, b, c, s1, s1s, s2, s2s, d1, d1s, d2, d2s):
        with self.ordered():
            self.check(lib().bbk_kernel_scale_2x2(self.h, count, a, b, c, _ptr(s1), s1s,
                                                  _ptr(s2), s2s, _ptr(d1), d1s, _ptr(d2), d2s),
                       "scale_2x2")

    def kernel_update_trailing(self, rows, cols, inner, left, ls, work, ws, target, ts):
        with self.ordered():
            self.check(lib().bbk_kernel_update_trailing(self.h, rows, cols, inner, _ptr(left), ls,
                                                        _ptr(work), ws, _ptr(target), ts),
                       "update_trailing")


class PinnedMatrix:
    """An (n, lda) row-major float64 matrix in pinned host memory
    (bbk_host_alloc), the storage bbk_dsytrf_ooc factors in place.
    ``array`` is a numpy view of it, valid until close()."""

    def __init__(self, n: int, lda: Optional[int] = None):
        import numpy as np
        self.n = n
        self.lda = n if lda is None else lda
        assert self.lda >= n > 0
        p = ctypes.c_void_p()
        rc = lib().bbk_host_alloc(ctypes.byref(p), 8 * n * self.lda)
        if rc != 0:
            raise BbkError(rc, "bbk_host_alloc failed")
        self.ptr = p.value
        buf = (ctypes.c_double * (n * self.lda)).from_address(self.ptr)
        self.array = np.ctypeslib.as_array(buf).reshape(n, self.lda)

    def close(self):
        if getattr(self, "ptr", None):
            self.array = None
            lib().bbk_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchGraph:
    """An instantiated HIP graph of one batched factorization."""

    def __init__(self, g, keep, handle):
        self.g = g
        self._keep = keep
        self._handle = handle

    def launch(self):
        with self._handle.ordered():
            rc = lib().bbk_graph_launch(self.g)
        if rc != 0:
            raise BbkError(rc, "bbk_graph_launch failed")

    def close(self):
        if getattr(self, "g", None):
            lib().bbk_graph_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default: Optional[Handle] = None


def default_handle() -> Handle:
    global _default
    if _default is None:
        _default = Handle()
    return _default


def padded_ld(n: int) -> int:
    """Leading dimension used for device copies: a multiple of 64 elements
    that is not a multiple of 1024 (avoids one-channel column walks)."""
    ld = (n + 63) // 64 * 64
    if ld % 1024 == 0:
        ld += 64
    return ld

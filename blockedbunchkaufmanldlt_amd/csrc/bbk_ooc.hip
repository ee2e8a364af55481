// bbk_ooc.hip -- the binary64 factorization of a matrix that stays in
// pinned host memory, with HBM as a bounded working area (the
// non-GPU-resident setting the reference cites, README.md:24-25; the
// reference itself is in-core only, README.md:155-157).
//
// The steps are the in-core factorization's (src/block_ldlt_decomposition.cc:
// 50-59: FactorPanel, then UpdateTrailingBlocks), in two phases:
//  - while the trailing matrix [s, n)^2 is larger than the budget, the
//    panel kernel works on the host copy in place (its pivot and candidate
//    rows are read across PCIe; every access stays inside [s, n)^2), the
//    interchanges of the rows above the panel are applied there too, and
//    the trailing update streams column slabs through HBM: per slab of
//    tile columns [ja, jb), rows [e, c1) x columns [c0, c1) are copied in,
//    updated by the MFMA tiles with the panel's U rows (copied once per
//    panel) and W rows, and copied back;
//  - from the first panel whose trailing matrix fits, that matrix is
//    copied to HBM once and the remaining panels run as in core on the
//    copy (a base shifted by its origin, so the kernels keep the matrix's
//    own indices); the rows above the copy take their interchanges in host
//    memory; the copy returns at the end.
// Every kernel is the in-core one, and every C element takes the same MFMA
// chain whichever slab holds it, so factors, perm and piv are bitwise those
// of bbk_dsytrf.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/bbk_ldlt.h"
#include "bbk_handle.hpp"

namespace bbk {
namespace {

// Row pitch of a device copy of m columns: a multiple of 64 elements,
// never a multiple of 1024 (a power-of-two pitch would put a column walk
// on one HBM channel).
int64_t copy_pitch(int64_t m) {
  int64_t l = (m + 63) / 64 * 64;
  if (l % 1024 == 0) l += 64;
  return std::max<int64_t>(l, 64);
}

size_t round256(size_t v) { return (v + 255) / 256 * 256; }

// The base that addresses a copy whose element (0, 0) is the matrix's
// (r0, c0) with the matrix's own indices: base[i ld + j] = copy[(i - r0) ld
// + (j - c0)].  Integer arithmetic: the shifted base itself is never
// dereferenced.
double* shifted(double* copy, int64_t ld, int64_t r0, int64_t c0) {
  return reinterpret_cast<double*>(reinterpret_cast<uintptr_t>(copy) -
                                   static_cast<uintptr_t>((r0 * ld + c0) * sizeof(double)));
}

// Device address of host memory from hipHostMalloc (bbk_host_alloc) that
// covers [p, p + bytes); NULL when it is not such memory.
double* pinned_device_ptr(double* p, size_t bytes) {
  hipPointerAttribute_t a0, a1;
  if (hipPointerGetAttributes(&a0, p) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  char* last = reinterpret_cast<char*>(p) + bytes - 1;
  if (hipPointerGetAttributes(&a1, last) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (a0.type != hipMemoryTypeHost || a1.type != hipMemoryTypeHost || !a0.devicePointer ||
      !a1.devicePointer)
    return nullptr;
  // one allocation: [p, p + bytes) inside the range that holds p
  void* start = nullptr;
  size_t size = 0;
  if (hipPointerGetAttribute(&start, HIP_POINTER_ATTRIBUTE_RANGE_START_ADDR, p) != hipSuccess ||
      hipPointerGetAttribute(&size, HIP_POINTER_ATTRIBUTE_RANGE_SIZE, p) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  const uintptr_t lo = reinterpret_cast<uintptr_t>(start);
  const uintptr_t at = reinterpret_cast<uintptr_t>(p);
  if (at < lo || at - lo > size || size - (at - lo) < bytes) return nullptr;
  return reinterpret_cast<double*>(a0.devicePointer);
}

// Interchange history: per panel, the count then the (a, r) pairs.
constexpr int kHistInts = 1 + 2 * (kPanelMaxWidth + 1);

__global__ void k_keep_swaps(const int* __restrict__ state, const int* __restrict__ swaps,
                             int* __restrict__ dst) {
  const int t = threadIdx.x;
  const int ns = state[PanelState::kNumSwaps];
  if (t == 0) dst[0] = ns;
  for (int i = t; i < 2 * ns; i += blockDim.x) dst[1 + i] = swaps[i];
}

// out(i, j) = in(i, src[j]) for rows < rows, columns < cols.
__global__ void k_gather_cols(const double* __restrict__ in, int64_t ldi,
                              double* __restrict__ out, int64_t ldo, const int* __restrict__ src,
                              int cols) {
  const int64_t i = blockIdx.y;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < cols; j += gridDim.x * blockDim.x)
    out[i * ldo + j] = in[i * ldi + src[j]];
}

// Everything one call allocates, released on every return path after the
// streams have drained.
struct OocResources {
  hipStream_t main = nullptr;
  void* work = nullptr;
  int* perm = nullptr;
  int* piv = nullptr;
  int* hist = nullptr;     // device interchange history
  int* src = nullptr;      // device column map of the final gathers
  int* state = nullptr;    // pinned host copy of the panel state
  hipStream_t up = nullptr, down = nullptr;
  hipEvent_t ready = nullptr, in[2] = {nullptr, nullptr}, done[2] = {nullptr, nullptr},
             out[2] = {nullptr, nullptr};
  hipError_t create() {
    hipError_t e = hipStreamCreateWithFlags(&up, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&down, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ready, hipEventDisableTiming);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
      e = hipEventCreateWithFlags(&in[i], hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&done[i], hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&out[i], hipEventDisableTiming);
    }
    return e;
  }
  ~OocResources() {
    if (main) (void)hipStreamSynchronize(main);
    if (up) (void)hipStreamSynchronize(up);
    if (down) (void)hipStreamSynchronize(down);
    for (hipEvent_t ev : {ready, in[0], in[1], done[0], done[1], out[0], out[1]})
      if (ev) (void)hipEventDestroy(ev);
    if (up) (void)hipStreamDestroy(up);
    if (down) (void)hipStreamDestroy(down);
    if (work) (void)hipFree(work);
    if (perm) (void)hipFree(perm);
    if (piv) (void)hipFree(piv);
    if (hist) (void)hipFree(hist);
    if (src) (void)hipFree(src);
    if (state) (void)hipHostFree(state);
  }
};

// Widest slab of tile columns (1024 columns): a slab copies its diagonal
// block whole, so narrow slabs keep the copied lower half small.
constexpr int kMaxSlabTiles = 16;

}  // namespace

int ooc_factorize(Handle* h, int n, int nb, double* A, int64_t lda, int* perm, int* piv,
                  size_t budget) {
  if (n <= 0) return -2;
  if (nb < 0) return -3;
  if (!A) return -4;
  if (lda < n) return -5;
  if (!perm) return -6;
  if (!piv) return -7;
  const size_t span = (static_cast<size_t>(n - 1) * lda + n) * sizeof(double);
  double* Ad = pinned_device_ptr(A, span);
  if (!Ad)
    return h->fail(-4, "bbk_dsytrf_ooc: A must lie in one pinned host allocation "
                       "(bbk_host_alloc) covering (n - 1) lda + n elements");
  int nt = 0, max_groups = 0;
  hipError_t e = hipSuccess;
  const int pc = panel_plan<double>(n, h->panel_threads, &nt, &max_groups, &e);
  if (pc == BBK_ERR_HIP) return h->hip_fail(e, "occupancy query");
  if (pc != BBK_OK)
    return h->fail(BBK_ERR_UNSUPPORTED,
                   "matrix order exceeds the co-resident panel grid (n <= 131072)");
  const int b = std::min(nb == 0 ? default_panel_width(n, sizeof(double)) : nb,
                         panel_max_width<double>(nt));
  constexpr int TM = 64;
  if (slab_tile_width<double>() != TM) return h->fail(BBK_ERR_UNSUPPORTED, "slab tile width");
  // Smallest working area: the U rows of one panel and two slab buffers of
  // one tile column of the full height.
  const size_t u_bytes = round256(sizeof(double) * (b + 1) * static_cast<size_t>(copy_pitch(n)));
  const size_t col_bytes = round256(sizeof(double) * static_cast<size_t>(n) * 64);
  const size_t min_bytes = u_bytes + 2 * col_bytes;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    e = hipMemGetInfo(&fr, &tot);
    if (e != hipSuccess) return h->hip_fail(e, "hipMemGetInfo");
    const size_t margin = size_t(1) << 30;
    budget = fr > 2 * margin ? fr - margin : fr / 2;
  }
  if (budget < min_bytes)
    return h->fail(BBK_ERR_UNSUPPORTED,
                   "bbk_dsytrf_ooc: device budget below (nb + 1) ldd + 128 n elements");
  const size_t full_bytes = sizeof(double) * static_cast<size_t>(n) * copy_pitch(n);
  const size_t area = std::min(budget, std::max(full_bytes, min_bytes));
  const int max_panels = (n + b - 1) / b + 1;

  FactorScratch fs;
  int rc = factor_setup<double>(h, n, &fs);
  if (rc != BBK_OK) return rc;
  hipStream_t st = h->stream;
  OocResources d;
  d.main = st;
  e = d.create();
  if (e == hipSuccess) e = hipMalloc(&d.work, area);
  if (e == hipSuccess) e = hipMalloc(&d.perm, sizeof(int) * n);
  if (e == hipSuccess) e = hipMalloc(&d.piv, sizeof(int) * n);
  if (e == hipSuccess) e = hipMalloc(&d.hist, sizeof(int) * kHistInts * max_panels);
  if (e == hipSuccess) e = hipMalloc(&d.src, sizeof(int) * n);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&d.state), 64);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return h->fail(BBK_ERR_ALLOC, "bbk_dsytrf_ooc: working area");
  }
  char* work = static_cast<char*>(d.work);
  const double* Wg = static_cast<const double*>(fs.Wg);
  launch_init(n, d.perm, d.piv, fs.state, h->d_status, st);

  auto fits = [&](int start) {
    const int64_t m = n - start;
    return sizeof(double) * static_cast<size_t>(m) * copy_pitch(m) <= area;
  };
  // Finished rows of the panels run on the host copy: their interchanges
  // by later panels are deferred to one gather per row at the end (no
  // kernel reads them before that).
  std::vector<int> host_s, host_e;
  int panels = 0, s_nom = 0, s_act = 0;
  // phase 1: the trailing matrix in host memory
  for (; s_nom < n && !fits(s_act); s_nom += b, ++panels) {
    // the panel reads what the previous panel's slabs wrote back
    e = hipStreamSynchronize(d.down);
    if (e != hipSuccess) return h->hip_fail(e, "slab copies");
    e = queue_panel<double>(n, b, s_nom, Ad, lda, d.perm, d.piv, fs, nt, max_groups, &h->epoch,
                            st, h->d_status, h->wait_ticks());
    if (e != hipSuccess) return h->hip_fail(e, "panel launch");
    hipLaunchKernelGGL(k_keep_swaps, dim3(1), dim3(256), 0, st, fs.state, fs.swaps,
                       d.hist + (size_t)panels * kHistInts);
    e = hipMemcpyAsync(d.state, fs.state, sizeof(int) * PanelState::kInts,
                       hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return h->hip_fail(e, "panel state read-back");
    const int s = d.state[PanelState::kStart];
    const int kk = d.state[PanelState::kWidth];
    const int pe = s + kk;
    if (kk <= 0) {
      // a panel wait failed (the status word says which); stop here
      rc = h->check_status();
      return rc != BBK_OK ? rc : h->fail(BBK_ERR_HIP, "bbk_dsytrf_ooc: empty panel");
    }
    host_s.push_back(s);
    host_e.push_back(pe);
    s_act = pe;
    if (pe >= n) continue;
    const int64_t m = n - pe;
    const int64_t ldu = copy_pitch(m);
    double* ub = reinterpret_cast<double*>(work);
    e = hipMemcpy2DAsync(ub, sizeof(double) * ldu, A + (int64_t)s * lda + pe,
                         sizeof(double) * lda, sizeof(double) * m, kk, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return h->hip_fail(e, "U rows to the device");
    e = hipEventRecord(d.ready, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(d.up, d.ready, 0);
    if (e != hipSuccess) return h->hip_fail(e, "slab ordering");
    const double* uv = shifted(ub, ldu, s, pe);
    const size_t u_used = round256(sizeof(double) * kk * ldu);
    const size_t buf_room = (area - u_used) / 2 / 256 * 256;
    double* cbuf[2] = {reinterpret_cast<double*>(work + u_used),
                       reinterpret_cast<double*>(work + u_used + buf_room)};
    const int tiles = static_cast<int>((m + TM - 1) / TM);
    // slabs alternate between two buffers: copy in on `up`, update on the
    // handle's stream, copy out on `down`, so the two copy directions and
    // the update overlap
    int slab = 0;
    for (int ja = 0; ja < tiles; ++slab) {
      const int bi = slab & 1;
      // the widest run of tile columns (at most kMaxSlabTiles) whose rows
      // [pe, c1) fit one buffer
      auto slab_bytes = [&](int j1) {
        const int64_t c0 = pe + (int64_t)ja * TM, c1 = std::min<int64_t>(n, pe + (int64_t)j1 * TM);
        return sizeof(double) * static_cast<size_t>(c1 - pe) * copy_pitch(c1 - c0);
      };
      int jb = ja + 1;
      while (jb < tiles && jb - ja < kMaxSlabTiles && slab_bytes(jb + 1) <= buf_room) ++jb;
      const int64_t c0 = pe + (int64_t)ja * TM, c1 = std::min<int64_t>(n, pe + (int64_t)jb * TM);
      const int64_t ldc = copy_pitch(c1 - c0);
      double* hsrc = A + (int64_t)pe * lda + c0;
      double* cb = cbuf[bi];
      if (slab >= 2) e = hipStreamWaitEvent(d.up, d.out[bi], 0);
      if (e == hipSuccess)
        e = hipMemcpy2DAsync(cb, sizeof(double) * ldc, hsrc, sizeof(double) * lda,
                             sizeof(double) * (c1 - c0), c1 - pe, hipMemcpyHostToDevice, d.up);
      if (e == hipSuccess) e = hipEventRecord(d.in[bi], d.up);
      if (e == hipSuccess) e = hipStreamWaitEvent(st, d.in[bi], 0);
      if (e != hipSuccess) return h->hip_fail(e, "slab to the device");
      launch_slab_update<double>(shifted(cb, ldc, pe, c0), ldc, uv, ldu, Wg, fs.ldw, fs.state, n,
                                 ja, jb, st);
      e = hipEventRecord(d.done[bi], st);
      if (e == hipSuccess) e = hipStreamWaitEvent(d.down, d.done[bi], 0);
      if (e == hipSuccess)
        e = hipMemcpy2DAsync(hsrc, sizeof(double) * lda, cb, sizeof(double) * ldc,
                             sizeof(double) * (c1 - c0), c1 - pe, hipMemcpyDeviceToHost, d.down);
      if (e == hipSuccess) e = hipEventRecord(d.out[bi], d.down);
      if (e != hipSuccess) return h->hip_fail(e, "slab to the host");
      ja = jb;
    }
  }
  e = hipStreamSynchronize(d.down);
  if (e != hipSuccess) return h->hip_fail(e, "slab copies");
  // phase 2: the trailing matrix [w0, n)^2 in HBM
  const int w0 = s_act;
  if (s_nom < n) {
    const int64_t m = n - w0;
    const int64_t ldwin = copy_pitch(m);
    double* win = reinterpret_cast<double*>(work);
    e = hipMemcpy2DAsync(win, sizeof(double) * ldwin, A + (int64_t)w0 * lda + w0,
                         sizeof(double) * lda, sizeof(double) * m, m, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return h->hip_fail(e, "trailing matrix to the device");
    double* av = shifted(win, ldwin, w0, w0);
    for (; s_nom < n; s_nom += b, ++panels) {
      e = queue_panel<double>(n, b, s_nom, av, ldwin, d.perm, d.piv, fs, nt, max_groups,
                              &h->epoch, st, h->d_status, h->wait_ticks());
      if (e != hipSuccess) return h->hip_fail(e, "panel launch");
      if (w0 > 0)
        hipLaunchKernelGGL(k_keep_swaps, dim3(1), dim3(256), 0, st, fs.state, fs.swaps,
                           d.hist + (size_t)panels * kHistInts);
      // interchanges of the finished rows inside the copy
      const int rows_bound = s_nom > 0 ? std::min(n, s_nom + panels) : 0;
      launch_trailing_update<double>(av, ldwin, Wg, fs.ldw, fs.state, fs.swaps, n,
                                     n - s_nom - b, rows_bound, st, w0);
    }
    e = hipMemcpy2DAsync(A + (int64_t)w0 * lda + w0, sizeof(double) * lda, win,
                         sizeof(double) * ldwin, sizeof(double) * m, m, hipMemcpyDeviceToHost,
                         st);
    if (e != hipSuccess) return h->hip_fail(e, "trailing matrix to the host");
  }
  if (!host_s.empty()) {
    // the deferred interchanges: panel P's rows [s_P, e_P) take the
    // composition of every later panel's interchanges, walked from the
    // last panel back: after panels Q.. are folded in, src(j) is the column
    // (as of the end of panel Q - 1) whose value ends at column j
    std::vector<int> hist(static_cast<size_t>(panels) * kHistInts);
    e = hipMemcpyAsync(hist.data(), d.hist, sizeof(int) * hist.size(), hipMemcpyDeviceToHost,
                       st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return h->hip_fail(e, "interchange history");
    std::vector<int> src(n), inv(n), xs, ys, js;
    for (int j = 0; j < n; ++j) src[j] = inv[j] = j;
    const int nh = static_cast<int>(host_s.size());
    for (int q = panels - 1; q >= 1; --q) {
      const int* rec = hist.data() + (size_t)q * kHistInts;
      const int ns = rec[0];
      // tau(x): the position before panel q's interchanges of the value at
      // x after them (its pairs undone in reverse order)
      xs.clear();
      for (int t = 0; t < 2 * ns; ++t) xs.push_back(rec[1 + t]);
      std::sort(xs.begin(), xs.end());
      xs.erase(std::unique(xs.begin(), xs.end()), xs.end());
      ys.clear();
      js.clear();
      for (int x : xs) {
        int y = x;
        for (int t = ns - 1; t >= 0; --t) {
          const int a = rec[1 + 2 * t], r = rec[2 + 2 * t];
          y = y == a ? r : y == r ? a : y;
        }
        js.push_back(inv[x]);
        ys.push_back(y);
      }
      for (size_t u = 0; u < js.size(); ++u) {
        src[js[u]] = ys[u];
        inv[ys[u]] = js[u];
      }
      const int P = q - 1;
      if (P >= nh) continue;
      const int ps = host_s[P], pe = host_e[P];
      int first = pe;
      while (first < n && src[first] == first) ++first;
      if (first >= n) continue;  // nothing moved past this panel
      // rows [ps, pe) x columns [first, n): copy in, gather, copy out
      const int cols = n - first;
      const int64_t ldx = copy_pitch(cols);
      const int rows_per = static_cast<int>(std::max<size_t>(
          1, std::min<size_t>(pe - ps, area / (2 * sizeof(double) * ldx))));
      std::vector<int> rel(cols);
      for (int j = 0; j < cols; ++j) rel[j] = src[first + j] - first;
      e = hipMemcpy(d.src, rel.data(), sizeof(int) * cols, hipMemcpyHostToDevice);
      if (e != hipSuccess) return h->hip_fail(e, "column map");
      double* xb = reinterpret_cast<double*>(work);
      double* yb = xb + (size_t)rows_per * ldx;
      for (int r0 = ps; r0 < pe; r0 += rows_per) {
        const int rn = std::min(rows_per, pe - r0);
        double* hrow = A + (int64_t)r0 * lda + first;
        e = hipMemcpy2DAsync(xb, sizeof(double) * ldx, hrow, sizeof(double) * lda,
                             sizeof(double) * cols, rn, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return h->hip_fail(e, "rows to the device");
        hipLaunchKernelGGL(k_gather_cols, dim3((cols + 255) / 256, rn), dim3(256), 0, st, xb,
                           ldx, yb, ldx, d.src, cols);
        e = hipMemcpy2DAsync(hrow, sizeof(double) * lda, yb, sizeof(double) * ldx,
                             sizeof(double) * cols, rn, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return h->hip_fail(e, "rows to the host");
      }
      e = hipStreamSynchronize(st);  // d.src is rewritten for the next panel
      if (e != hipSuccess) return h->hip_fail(e, "deferred interchanges");
    }
  }
  launch_factor_info<double>(n, Ad, lda, d.piv, h->d_status, st);
  e = hipGetLastError();
  if (e != hipSuccess) return h->hip_fail(e, "kernel launch");
  e = hipMemcpyAsync(perm, d.perm, sizeof(int) * n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(piv, d.piv, sizeof(int) * n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return h->hip_fail(e, "bbk_dsytrf_ooc");
  return h->check_status();
}

}  // namespace bbk

extern "C" {

int bbk_dsytrf_ooc(bbk_handle_t handle, int n, int nb, double* A, int64_t lda, int* perm,
                   int* piv, size_t device_bytes) {
  if (!handle) return -1;
  return bbk::ooc_factorize(reinterpret_cast<bbk::Handle*>(handle), n, nb, A, lda, perm, piv,
                            device_bytes);
}

int bbk_host_alloc(void** p, size_t bytes) {
  if (!p) return -1;
  *p = nullptr;
  if (bytes == 0) return -2;
  hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return BBK_ERR_ALLOC;
  }
  return BBK_OK;
}

int bbk_host_free(void* p) {
  if (!p) return BBK_OK;
  return hipHostFree(p) == hipSuccess ? BBK_OK : BBK_ERR_HIP;
}

}  // extern "C"

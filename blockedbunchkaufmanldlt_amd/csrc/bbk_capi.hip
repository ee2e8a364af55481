// bbk_capi.hip -- the C ABI (include/bbk_ldlt.h): handle, workspace and
// the host-side panel loop that drives the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bbk_ldlt.h"
#include "bbk_handle.hpp"
#include "bbk_kernels.hpp"

using bbk::PanelArgs;

using bbk::PanelState;

namespace bbk {

int Handle::fail(int code, const char* what) {
  last_error = what;
  return code;
}

int Handle::hip_fail(hipError_t e, const char* what) {
  last_error = std::string(what) + ": " + hipGetErrorString(e);
  return BBK_ERR_HIP;
}

// Device scratch grows monotonically and is reused across calls.
int Handle::reserve(size_t bytes) {
  if (bytes <= ws_bytes) return BBK_OK;
  if (ws) {
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    hipFree(ws);
    ws = nullptr;
    ws_bytes = 0;
  }
  hipError_t e = hipMalloc(&ws, bytes);
  if (e != hipSuccess) {
    ws = nullptr;
    return hip_fail(e, "hipMalloc(workspace)");
  }
  ws_bytes = bytes;
  ++ws_gen;
  return BBK_OK;
}

int Handle::reserve_buf(void** p, size_t* cur, size_t want, const char* what, unsigned* gen) {
  if (want <= *cur) return BBK_OK;
  if (*p) {
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    (void)hipFree(*p);
    *p = nullptr;
    *cur = 0;
  }
  hipError_t e = hipMalloc(p, want);
  if (e != hipSuccess) {
    *p = nullptr;
    return hip_fail(e, what);
  }
  *cur = want;
  if (gen) ++*gen;
  return BBK_OK;
}


int default_panel_width(int n, int elem_bytes) {
  (void)elem_bytes;
  if (n <= 256) return 32;
  if (n <= 1024) return 64;
  return 128;
}

// The panel variant for order n: workgroups of kPanelThreads columns while
// the grid of ceil(n / 256) groups fits co-resident (n <= 65536), else of
// kPanelThreadsWide columns (n <= 131072); force (256 or 512) allows that
// variant only.  *max_groups: co-resident workgroups of that variant.
// BBK_ERR_UNSUPPORTED past both.
template <typename T>
int panel_plan(int n, int force, int* nt, int* max_groups, hipError_t* err) {
  *err = hipSuccess;
  for (int cand : {kPanelThreads, kPanelThreadsWide}) {
    if (force != 0 && cand != force) continue;
    int g = 0;
    hipError_t e = panel_max_groups<T>(cand, &g);
    if (e != hipSuccess) {
      *err = e;
      return BBK_ERR_HIP;
    }
    if ((n + cand - 1) / cand <= std::min(g, kMaxPanelGroupsLds)) {
      *nt = cand;
      *max_groups = g;
      return BBK_OK;
    }
  }
  return BBK_ERR_UNSUPPORTED;
}

template int panel_plan<double>(int, int, int*, int*, hipError_t*);
template int panel_plan<float>(int, int, int*, int*, hipError_t*);

// The panel loop shared by every factorization entry point: queues the
// init kernel, then per panel the cooperative panel kernel and the trailing
// update (with the interchanges of the rows above the panel) on `st`.
// fs: the carved scratch; *epoch: the next exchange-tag range (advanced by
// 1024 per launch).  h (may be NULL) receives timing marks and the panel
// count.
template <typename T>
hipError_t queue_panel(int n, int b, int s_nom, T* A, int64_t lda, int* perm, int* piv,
                       const FactorScratch& fs, int nt, int max_groups, int* epoch,
                       hipStream_t st, unsigned* status, long long wait_ticks) {
  const int groups = (n - s_nom + nt - 1) / nt;
  PanelArgs pa;
  pa.n = n;
  pa.b = b;
  pa.lda = lda;
  pa.state = fs.state;
  pa.swaps = fs.swaps;
  pa.perm = perm;
  pa.piv = piv;
  pa.xrec = fs.xrec;
  pa.epoch = *epoch;
  *epoch += 1024;
  pa.Wg = fs.Wg;
  pa.ldw = fs.ldw;
  // Small panels run on one XCD: the grid is spread over the 8 XCDs and
  // only every 8th workgroup works, when all of them fit co-resident.
  pa.xcd_stride = 8 * groups <= max_groups ? 8 : 1;
  pa.status = status;
  pa.wait_ticks = wait_ticks;
  return launch_panel<T>(A, pa, groups * pa.xcd_stride, nt, st);
}

template hipError_t queue_panel<double>(int, int, int, double*, int64_t, int*, int*,
                                        const FactorScratch&, int, int, int*, hipStream_t,
                                        unsigned*, long long);
template hipError_t queue_panel<float>(int, int, int, float*, int64_t, int*, int*,
                                       const FactorScratch&, int, int, int*, hipStream_t,
                                       unsigned*, long long);

template <typename T>
int factor_loop(int n, int b, T* A, int64_t lda, int* perm, int* piv, const FactorScratch& fs,
                int nt, int max_groups, int* epoch, hipStream_t st, Handle* h, unsigned* status,
                long long wait_ticks) {
  b = std::min(b, panel_max_width<T>(nt));
  launch_init(n, perm, piv, fs.state, status, st);
  int panels = 0;
  for (int s_nom = 0; s_nom < n; s_nom += b, ++panels) {
    if (h) h->timing_mark(Handle::kPanel, true);
    hipError_t e = queue_panel<T>(n, b, s_nom, A, lda, perm, piv, fs, nt, max_groups, epoch, st,
                                  status, wait_ticks);
    if (e != hipSuccess) return h ? h->hip_fail(e, "panel launch") : BBK_ERR_HIP;
    if (h) h->timing_mark(Handle::kPanel, false);
    // Rows above this panel: at most s_nom + panels (each earlier panel
    // can widen by one column); their interchanges run in the same launch
    // as the trailing update.
    const int rows_bound = s_nom > 0 ? std::min(n, s_nom + panels) : 0;
    if (h) h->timing_mark(Handle::kUpdate, true);
    launch_trailing_update<T>(A, lda, reinterpret_cast<const T*>(fs.Wg), fs.ldw, fs.state,
                              fs.swaps, n, n - s_nom - b, rows_bound, st);
    if (h) h->timing_mark(Handle::kUpdate, false);
  }
  if (h && h->timing) h->panels += panels;
  if (status) launch_factor_info<T>(n, A, lda, piv, status, st);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return h ? h->hip_fail(e, "kernel launch") : BBK_ERR_HIP;
  return BBK_OK;
}

template int factor_loop<double>(int, int, double*, int64_t, int*, int*, const FactorScratch&,
                                 int, int, int*, hipStream_t, Handle*, unsigned*, long long);
template int factor_loop<float>(int, int, float*, int64_t, int*, int*, const FactorScratch&,
                                int, int, int*, hipStream_t, Handle*, unsigned*, long long);

template <typename T>
int factorize(Handle* h, int n, int nb, T* A, int64_t lda, int* perm,
              int* piv) {
  if (n <= 0) return -2;
  if (nb < 0) return -3;
  if (!A) return -4;
  if (lda < n) return -5;
  if (!perm) return -6;
  if (!piv) return -7;
  int b = nb == 0 ? default_panel_width(n, sizeof(T)) : nb;
  int nt = 0, max_groups = 0;
  hipError_t e = hipSuccess;
  const int pc = panel_plan<T>(n, h->panel_threads, &nt, &max_groups, &e);
  if (pc == BBK_ERR_HIP) return h->hip_fail(e, "occupancy query");
  if (pc != BBK_OK)
    return h->fail(BBK_ERR_UNSUPPORTED,
                   "matrix order exceeds the co-resident panel grid (n <= 131072)");
  FactorScratch fs;
  int rc = factor_setup<T>(h, n, &fs);
  if (rc != BBK_OK) return rc;
  return factor_loop<T>(n, b, A, lda, perm, piv, fs, nt, max_groups, &h->epoch, h->stream, h,
                        h->d_status, h->wait_ticks());
}

template <typename T>
int factor_setup(Handle* h, int n, FactorScratch* fs) {
  const size_t bytes = factor_scratch_bytes<T>(n, nullptr, nullptr);
  int rc = h->reserve(bytes);
  if (rc != BBK_OK) return rc;
  factor_scratch_bytes<T>(n, fs, reinterpret_cast<char*>(h->ws));
  // Exchange-record tags must never repeat within the live record area:
  // each launch takes a fresh range of 1024 tags; a (re)allocated area, an
  // area laid out for another order or type, or a wrapped tag range is
  // cleared first (fill kernel).
  const long layout = static_cast<long>(n) * 16 + static_cast<long>(sizeof(T));
  if (h->ws_gen != h->ws_clear_gen || h->ws_layout != layout || h->epoch > (1 << 30)) {
    launch_zero_u32(reinterpret_cast<unsigned*>(fs->xrec),
                    static_cast<int>(fs->xrec_bytes / sizeof(unsigned)), h->stream);
    h->ws_clear_gen = h->ws_gen;
    h->ws_layout = layout;
    h->epoch = 1024;
  }
  return BBK_OK;
}

template int factor_setup<double>(Handle*, int, FactorScratch*);
template int factor_setup<float>(Handle*, int, FactorScratch*);

template int factorize<double>(Handle*, int, int, double*, int64_t, int*, int*);
template int factorize<float>(Handle*, int, int, float*, int64_t, int*, int*);

int Handle::check_status() {
  if (!d_status) return BBK_OK;
  unsigned w[kStatusWords];
  hipError_t e = hipMemcpyAsync(w, d_status, sizeof(w), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return hip_fail(e, "read device status");
  if (w[kStatusError] != 0) {
    launch_zero_u32(d_status + kStatusError, 1, stream);
    (void)hipStreamSynchronize(stream);
    last_error = std::string("a device wait exceeded its time limit (") +
                 ((w[kStatusError] & kErrPanelWait) ? "panel exchange" : "") +
                 ((w[kStatusError] & kErrPanelWait) && (w[kStatusError] & kErrSweepWait) ? ", " : "") +
                 ((w[kStatusError] & kErrSweepWait) ? "triangular sweep" : "") +
                 "); the results of the work queued since the last check are invalid";
    return BBK_ERR_HIP;
  }
  return BBK_OK;
}

// ---- optional event timing of the kernel groups -------------------------
void Handle::timing_reset() {
  for (auto& v : marks) {
    for (hipEvent_t ev : v) hipEventDestroy(ev);
    v.clear();
  }
}

void Handle::timing_mark(int group, bool begin) {
  if (!timing) return;
  hipEvent_t ev;
  if (hipEventCreate(&ev) != hipSuccess) return;
  hipEventRecord(ev, stream);
  marks[group].push_back(ev);
  (void)begin;
}

double Handle::timing_total(int group) {
  double total = 0.0;
  std::vector<hipEvent_t>& v = marks[group];
  for (size_t i = 0; i + 1 < v.size(); i += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, v[i], v[i + 1]) == hipSuccess) total += ms;
  }
  return total;
}

Handle::~Handle() {
  for (auto& v : marks)
    for (hipEvent_t ev : v) hipEventDestroy(ev);
  if (ws) (void)hipFree(ws);
  if (solve_ws) (void)hipFree(solve_ws);
  if (berr_ws) (void)hipFree(berr_ws);
  if (mixed_ws) (void)hipFree(mixed_ws);
  if (batch.buf) (void)hipFree(batch.buf);
  if (batch.ws) (void)hipFree(batch.ws);
  if (d_status) (void)hipFree(d_status);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

}  // namespace bbk

using bbk::Handle;

extern "C" {

int bbk_create(bbk_handle_t* handle, void* stream) {
  if (!handle) return -1;
  Handle* h = new Handle();
  hipError_t e = hipGetDevice(&h->device);
  if (e != hipSuccess) {
    delete h;
    return BBK_ERR_HIP;
  }
  if (stream) {
    h->stream = reinterpret_cast<hipStream_t>(stream);
    h->own_stream = false;
  } else {
    // non-blocking: cooperative launches are refused on streams that
    // synchronize with the null stream
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete h;
      return BBK_ERR_HIP;
    }
    h->own_stream = true;
  }
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, h->device) == hipSuccess &&
      khz > 0)
    h->clock_khz = khz;
  e = hipMalloc(&h->d_status, sizeof(unsigned) * bbk::kStatusWords);
  if (e != hipSuccess) {
    h->d_status = nullptr;
    delete h;
    return BBK_ERR_ALLOC;
  }
  const unsigned init[bbk::kStatusWords] = {0u, bbk::kInfoNone, 0u, 0u};
  e = hipMemcpyAsync(h->d_status, init, sizeof(init), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    delete h;
    return BBK_ERR_HIP;
  }
  *handle = reinterpret_cast<bbk_handle_t>(h);
  return BBK_OK;
}

int bbk_get_stream(bbk_handle_t handle, void** stream) {
  if (!handle) return -1;
  if (!stream) return -2;
  *stream = reinterpret_cast<Handle*>(handle)->stream;
  return BBK_OK;
}

int bbk_destroy(bbk_handle_t handle) {
  if (!handle) return -1;
  Handle* h = reinterpret_cast<Handle*>(handle);
  hipStreamSynchronize(h->stream);
  delete h;
  return BBK_OK;
}

int bbk_synchronize(bbk_handle_t handle) {
  if (!handle) return -1;
  Handle* h = reinterpret_cast<Handle*>(handle);
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return h->hip_fail(e, "hipStreamSynchronize");
  return h->check_status();
}

int bbk_set_panel_threads(bbk_handle_t handle, int threads) {
  if (!handle) return -1;
  if (threads != 0 && threads != bbk::kPanelThreads && threads != bbk::kPanelThreadsWide)
    return -2;
  reinterpret_cast<Handle*>(handle)->panel_threads = threads;
  return BBK_OK;
}

int bbk_set_wait_limit(bbk_handle_t handle, int ms) {
  if (!handle) return -1;
  reinterpret_cast<Handle*>(handle)->wait_ms = ms;
  return BBK_OK;
}

int bbk_get_info(bbk_handle_t handle, int* info, int* nonfinite) {
  if (!handle) return -1;
  Handle* h = reinterpret_cast<Handle*>(handle);
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return h->hip_fail(e, "hipStreamSynchronize");
  unsigned w[bbk::kStatusWords];
  e = hipMemcpy(w, h->d_status, sizeof(w), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return h->hip_fail(e, "read device status");
  if (info) *info = w[bbk::kStatusInfo] == bbk::kInfoNone ? 0 : static_cast<int>(w[bbk::kStatusInfo]);
  if (nonfinite) *nonfinite = static_cast<int>(w[bbk::kStatusNonFinite]);
  return h->check_status();
}

const char* bbk_last_error(bbk_handle_t handle) {
  if (!handle) return "null handle";
  return reinterpret_cast<Handle*>(handle)->last_error.c_str();
}

int bbk_dsytrf(bbk_handle_t handle, int n, int nb, double* d_A, int64_t lda,
               int* d_perm, int* d_piv) {
  if (!handle) return -1;
  return bbk::factorize<double>(reinterpret_cast<Handle*>(handle), n, nb, d_A,
                                lda, d_perm, d_piv);
}

int bbk_ssytrf(bbk_handle_t handle, int n, int nb, float* d_A, int64_t lda,
               int* d_perm, int* d_piv) {
  if (!handle) return -1;
  return bbk::factorize<float>(reinterpret_cast<Handle*>(handle), n, nb, d_A,
                               lda, d_perm, d_piv);
}

int bbk_dsytrf_host(bbk_handle_t handle, int n, int nb, double* A,
                    int64_t lda, int* perm, int* piv) {
  if (!handle) return -1;
  Handle* h = reinterpret_cast<Handle*>(handle);
  if (n <= 0) return -2;
  if (nb < 0) return -3;
  if (!A) return -4;
  if (lda < n) return -5;
  if (!perm) return -6;
  if (!piv) return -7;
  // Device copy with a padded leading dimension (a power-of-two row pitch
  // would put a column walk on one HBM channel).
  int64_t ldd = (n + 63) / 64 * 64;
  if (ldd % 1024 == 0) ldd += 64;
  double* dA = nullptr;
  int* dperm = nullptr;
  int* dpiv = nullptr;
  hipError_t e = hipMalloc(&dA, sizeof(double) * static_cast<size_t>(ldd) * n);
  if (e == hipSuccess) e = hipMalloc(&dperm, sizeof(int) * n);
  if (e == hipSuccess) e = hipMalloc(&dpiv, sizeof(int) * n);
  int rc = BBK_OK;
  if (e != hipSuccess) {
    rc = h->hip_fail(e, "hipMalloc");
  } else {
    e = hipMemcpy2DAsync(dA, sizeof(double) * ldd, A, sizeof(double) * lda,
                         sizeof(double) * n, n, hipMemcpyHostToDevice,
                         h->stream);
    if (e != hipSuccess) rc = h->hip_fail(e, "hipMemcpy2D H2D");
    if (rc == BBK_OK) rc = bbk::factorize<double>(h, n, nb, dA, ldd, dperm, dpiv);
    if (rc == BBK_OK) {
      e = hipMemcpy2DAsync(A, sizeof(double) * lda, dA, sizeof(double) * ldd,
                           sizeof(double) * n, n, hipMemcpyDeviceToHost,
                           h->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(perm, dperm, sizeof(int) * n, hipMemcpyDeviceToHost,
                           h->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(piv, dpiv, sizeof(int) * n, hipMemcpyDeviceToHost,
                           h->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
      if (e != hipSuccess) rc = h->hip_fail(e, "D2H copy");
      if (rc == BBK_OK) rc = h->check_status();
    }
  }
  hipStreamSynchronize(h->stream);
  if (dA) hipFree(dA);
  if (dperm) hipFree(dperm);
  if (dpiv) hipFree(dpiv);
  return rc;
}

int bbk_set_timing(bbk_handle_t handle, int enabled) {
  if (!handle) return -1;
  Handle* h = reinterpret_cast<Handle*>(handle);
  hipStreamSynchronize(h->stream);
  h->timing_reset();
  h->panels = 0;
  h->timing = enabled != 0;
  return BBK_OK;
}

int bbk_last_timing(bbk_handle_t handle, double* panel_ms, double* update_ms,
                    double* swap_ms, int* panels) {
  if (!handle) return -1;
  Handle* h = reinterpret_cast<Handle*>(handle);
  if (panel_ms) *panel_ms = h->timing_total(Handle::kPanel);
  if (update_ms) *update_ms = h->timing_total(Handle::kUpdate);
  if (swap_ms) *swap_ms = h->timing_total(Handle::kSwap);
  if (panels) *panels = h->panels;
  return BBK_OK;
}

}  // extern "C"

// bbk_handle.hpp -- state behind a bbk_handle_t.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <vector>

#include "bbk_batched.hpp"
#include "bbk_kernels.hpp"

namespace bbk {

struct Handle {
  enum { kPanel = 0, kUpdate = 1, kSwap = 2, kGroups = 3 };

  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // Device scratch regions, each grown on demand and kept for reuse.
  void* ws = nullptr;        // factorization: exchange area, W rows
  size_t ws_bytes = 0;
  // Record areas are cleared whenever their buffer was (re)allocated or
  // their layout (order, element size) changed since the last clearing:
  // an allocation can return the same address with unrelated contents.
  unsigned ws_gen = 0;             // bumped by every (re)allocation of ws
  unsigned ws_clear_gen = ~0u;     // ws_gen when the exchange area was cleared
  long ws_layout = -1;             // layout the exchange area was cleared for
  void* solve_ws = nullptr;  // triangular sweeps: flags, Z, V
  size_t solve_bytes = 0;
  unsigned solve_gen = 0;
  unsigned solve_clear_gen = ~0u;
  long solve_layout = -1;
  int solve_tag = 0;             // last record tag used by a sweep launch
  void* berr_ws = nullptr;   // residual partial sums and norms
  size_t berr_bytes = 0;
  void* mixed_ws = nullptr;  // refinement driver: factor copy, r, d
  size_t mixed_bytes = 0;
  double* last_anorm_dev = nullptr;
  // Device status words (kStatusWords unsigned, bbk_internal.hpp): wait
  // errors, the last factorization's singular-pivot info and non-finite flag.
  unsigned* d_status = nullptr;
  int wait_ms = 0;          // exchange / sweep wait limit: 0 = kDefaultWaitMs, < 0 = fail at once
  long long clock_khz = 100000;  // device wall-clock rate (ticks per ms)
  long long wait_ticks() const {
    if (wait_ms < 0) return -1;
    return static_cast<long long>(wait_ms == 0 ? kDefaultWaitMs : wait_ms) * clock_khz;
  }
  // Reads the status words after the stream has finished; a wait error is
  // cleared and returned as BBK_ERR_HIP (with last_error set).
  int check_status();
  int epoch = 1024;          // next panel-exchange tag range
  std::string last_error;
  bool timing = false;
  int panels = 0;
  std::vector<hipEvent_t> marks[kGroups];
  BatchPlan batch;                 // last prepared batched schedule
  int batch_grid_limit = 0;        // > 0: cap on persistent batched workgroups
  int panel_threads = 0;           // panel variant: 0 = by order, else 256 / 512
  BatchWidths batch_widths[2] = {kDefaultBatchWidthsF64, kDefaultBatchWidthsF32};

  ~Handle();
  int fail(int code, const char* what);
  int hip_fail(hipError_t e, const char* what);
  int reserve(size_t bytes);
  int reserve_buf(void** p, size_t* cur, size_t want, const char* what,
                  unsigned* gen = nullptr);
  int reserve_solve(size_t b) {
    return reserve_buf(&solve_ws, &solve_bytes, b, "solve", &solve_gen);
  }
  int reserve_berr(size_t b) { return reserve_buf(&berr_ws, &berr_bytes, b, "residual"); }
  int reserve_mixed(size_t b) { return reserve_buf(&mixed_ws, &mixed_bytes, b, "refinement"); }
  int next_solve_tag() { return ++solve_tag; }
  void timing_reset();
  void timing_mark(int group, bool begin);
  double timing_total(int group);
};

int default_panel_width(int n, int elem_bytes);

template <typename T>
int factorize(Handle* h, int n, int nb, T* A, int64_t lda, int* perm,
              int* piv);
// Scratch of a factorization of order n (exchange records cleared when
// their area or layout changed), queued on h->stream.
template <typename T>
int factor_setup(Handle* h, int n, FactorScratch* fs);
// One panel launch at nominal start s_nom (the actual start is in the
// panel state), taking the next exchange-tag range from *epoch.
template <typename T>
hipError_t queue_panel(int n, int b, int s_nom, T* A, int64_t lda, int* perm, int* piv,
                       const FactorScratch& fs, int nt, int max_groups, int* epoch,
                       hipStream_t st, unsigned* status, long long wait_ticks);
// The factorization's launch sequence on stream st with carved scratch fs
// (h: timing and error text, may be NULL).
template <typename T>
int factor_loop(int n, int b, T* A, int64_t lda, int* perm, int* piv, const FactorScratch& fs,
                int nt, int max_groups, int* epoch, hipStream_t st, Handle* h, unsigned* status,
                long long wait_ticks);
// The panel variant for order n (threads per workgroup nt, its co-resident
// workgroups max_groups; force: 0 = the smallest variant whose grid fits,
// else that variant only): BBK_OK, BBK_ERR_HIP (err set) or
// BBK_ERR_UNSUPPORTED when no allowed variant's grid fits.
template <typename T>
int panel_plan(int n, int force, int* nt, int* max_groups, hipError_t* err);
// Li: inverses of L's 128 x 128 diagonal blocks (binary32), used by the
// refinement's solves instead of substitution, with F's strictly lower
// triangle holding the mirror of the upper one; NULL substitutes (F upper
// only).  stop: device flag; when set, the sweeps return at once (NULL:
// always run).
template <typename T>
int solve(Handle* h, int n, int nrhs, const T* F, int64_t lda, const int* perm,
          const int* piv, T* B, int64_t ldb, const T* Li = nullptr,
          const int* stop = nullptr);
int batched_prepare(Handle* h, int count, const int* h_n, int elem_bytes);
template <typename T>
int batched_sytrf(Handle* h, int count, const int* h_n, T* const* A, const int64_t* lda,
                  int* const* perm, int* const* piv);
// stop: as for solve (the residual kernels return at once when it is set).
int berr(Handle* h, int n, int nrhs, const double* A, int64_t lda,
         const double* B, int64_t ldb, const double* X, int64_t ldx, double* R,
         int64_t ldr, double* d_eta, double anorm, bool want_norm,
         const int* stop = nullptr);

}  // namespace bbk

// bbk_kernels.hpp -- kernel launch interface between the C-ABI layer
// (bbk_capi.hip) and the kernel translation units.
#pragma once

#include "bbk_internal.hpp"

namespace bbk {

struct PanelArgs {
  int n;
  int b;            // target panel width (<= kPanelMaxWidth)
  int64_t lda;
  int* state;       // PanelState ints
  int* swaps;       // 2 * (kPanelMaxWidth + 1) ints: interchanges (a, r)
  int* perm;
  int* piv;
  XRec* xrec;       // panel_exchange_records(G, sizeof(T)) records
  int epoch;        // record tags of this launch are epoch + 1, epoch + 2, ...
  void* Wg;         // (kPanelMaxWidth + 1) x ldw elements of T
  int64_t ldw;
  int xcd_stride;   // 8: only workgroups blockIdx % 8 == 0 work (one XCD); 1: all
  unsigned* status; // kStatusWords device status words (may be NULL)
  long long wait_ticks;  // exchange waits: device-clock ticks before giving up (< 0: fail at once)
};

// Exchange records the panel kernel needs for a grid of `groups`.
size_t panel_exchange_records(int groups, int elem_bytes);

// nt: threads per workgroup of the panel variant (kPanelThreads or
// kPanelThreadsWide).
template <typename T>
hipError_t launch_panel(T* A, const PanelArgs& args, int groups, int nt,
                        hipStream_t stream);
template <typename T>
hipError_t panel_max_groups(int nt, int* groups);
// Trailing update of the panel just factored plus the interchange of the
// rows [rows_lo, min(rows_bound, panel start)) above it (rows_bound: upper
// bound of the panel start; rows_bound <= rows_lo: none).
template <typename T>
void launch_trailing_update(T* A, int64_t lda, const T* Wg, int64_t ldw,
                            const int* state, const int* swaps, int n, int m_bound,
                            int rows_bound, hipStream_t stream, int rows_lo = 0);
// The trailing update's tiles in tile columns [ja, jb) (all tiles I <= J of
// those columns), on copies: C(i, j) = C[i ldc + j] and U(s + q, i) =
// U[(s + q) ldu + i] with bases shifted by the copies' origins.
template <typename T>
int slab_tile_width();
template <typename T>
void launch_slab_update(T* C, int64_t ldc, const T* U, int64_t ldu, const T* Wg, int64_t ldw,
                        const int* state, int n, int ja, int jb, hipStream_t stream);
// After the last panel: the first singular pivot and the non-finite flag
// of the factorization into the status words.
template <typename T>
void launch_factor_info(int n, const T* A, int64_t lda, const int* piv, unsigned* status,
                        hipStream_t stream);
// perm = identity, piv = 0, panel state cleared; with status, the info and
// non-finite words of a new factorization reset.
void launch_init(int n, int* perm, int* piv, int* state, unsigned* status, hipStream_t stream);

// Scratch carve-up for one factorization of order n in type T: exchange
// records, panel state, the interchange list and the flushed W rows.  A
// carve for order n serves every order <= n.
struct FactorScratch {
  XRec* xrec;
  size_t xrec_bytes;
  int* state;
  int* swaps;
  void* Wg;
  int64_t ldw;
};

template <typename T>
inline size_t factor_scratch_bytes(int n, FactorScratch* fs, char* base) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = (off + bytes + 255) / 256 * 256;
    return base ? base + o : nullptr;
  };
  const int64_t ldw = (static_cast<int64_t>(n) + 63) / 64 * 64;
  const int groups = (n + kPanelThreads - 1) / kPanelThreads;
  FactorScratch f;
  f.xrec_bytes = sizeof(XRec) * panel_exchange_records(groups, sizeof(T));
  f.xrec = reinterpret_cast<XRec*>(take(f.xrec_bytes));
  f.state = reinterpret_cast<int*>(take(sizeof(int) * PanelState::kInts));
  f.swaps = reinterpret_cast<int*>(take(sizeof(int) * 2 * (kPanelMaxWidth + 1)));
  f.Wg = take(sizeof(T) * (kPanelMaxWidth + 1) * static_cast<size_t>(ldw));
  f.ldw = ldw;
  if (fs) *fs = f;
  return off;
}

void launch_zero_u32(unsigned* p, int count, hipStream_t stream);

// Launch of a kernel whose workgroups wait on each other's records (panel
// exchanges, triangular sweeps).  Its grid never exceeds the co-resident
// capacity (occupancy x CUs).  The cooperative launch guarantees that every
// workgroup is resident.  The runtime refuses a cooperative launch in a
// second process once another process on the device holds the wave-sync
// resource; with bounded = true such a launch falls back to an ordinary one
// (on a device not busy with other work its workgroups are all resident).
// That is safe only because every wait of these kernels is bounded: if a
// workgroup were not scheduled, the waits end at the time limit and the
// call reports BBK_ERR_HIP from the status word instead of hanging.
// Callers without a status word pass bounded = false and keep the
// cooperative launch only.
inline hipError_t launch_resident(const void* fn, dim3 grid, dim3 block, void** args,
                                  hipStream_t stream, bool bounded) {
  hipError_t e = hipLaunchCooperativeKernel(fn, grid, block, args, 0, stream);
  if (e == hipSuccess || !bounded) return e;
  (void)hipGetLastError();  // the refused launch's error is not sticky
  return hipLaunchKernel(fn, grid, block, args, 0, stream);
}

}  // namespace bbk

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
};

// Exchange records the panel kernel needs for a grid of `groups`.
size_t panel_exchange_records(int groups, int elem_bytes);

template <typename T>
hipError_t launch_panel(T* A, const PanelArgs& args, int groups,
                        hipStream_t stream);
template <typename T>
hipError_t panel_max_groups(int* groups);
// Trailing update of the panel just factored plus the interchange of the
// rows above it (rows_bound: upper bound of the panel start; 0 = none).
template <typename T>
void launch_trailing_update(T* A, int64_t lda, const T* Wg, int64_t ldw,
                            const int* state, const int* swaps, int n, int m_bound,
                            int rows_bound, hipStream_t stream);
void launch_init(int n, int* perm, int* piv, int* state, hipStream_t stream);
void launch_zero_u32(unsigned* p, int count, hipStream_t stream);

}  // namespace bbk

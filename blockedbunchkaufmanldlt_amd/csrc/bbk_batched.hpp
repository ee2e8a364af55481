// bbk_batched.hpp -- batched factorization: size classes, panel-width
// table and the per-handle schedule.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "bbk_internal.hpp"

namespace bbk {

constexpr int kBatchClasses = 8;    // n <= 32, 64, 128, 256, 512, 1024, 2048, larger
constexpr int kBatchMaxWidth = 64;  // widest panel a batched matrix may use
// Dynamic LDS per workgroup for the panel rows: six one-wavefront teams or
// two four-wavefront teams fit a CU's 160 KB.
constexpr int kBatchLdsBytesSmall = 24 * 1024;
constexpr int kBatchLdsBytes = 76 * 1024;
// Batches whose largest matrix exceeds this run in lockstep: per panel step
// one launch factors every matrix's panel (one workgroup each), a second
// applies all trailing updates as MFMA tiles spread over the whole GPU.
constexpr int kBatchLockstepN = 512;
// Batches whose largest matrix is at most this run resident: one
// workgroup per matrix keeps the matrix (packed upper triangle), its panel
// rows and its permutation in LDS for the whole factorization and applies
// the trailing updates with MFMA tiles on LDS data.
constexpr int kBatchResidentN = 256;
// Dynamic LDS of the resident kernel (one workgroup per CU).
constexpr int kBatchResidentLdsBytes = 152 * 1024;
// Team-mode code of the symmetric-workspace kernel (classes up to
// kBatchResidentN): one wavefront per matrix, a full symmetric copy of the
// matrix in a per-workgroup workspace slice, interchanges as index-map
// updates.
constexpr int kBatchTeamSym = 1;
__host__ __device__ constexpr int batch_lds_bytes(int nt) {
  return nt == 64 ? kBatchLdsBytesSmall : kBatchLdsBytes;
}

// Upper size bound of class c (the last class is unbounded).
__host__ __device__ constexpr int batch_class_bound(int c) {
  return c < kBatchClasses - 1 ? (32 << c) : 0x7fffffff;
}

// Per size class: panel width w, team nt and trailing-update tile ts.
//   nt  64 or 256 threads per matrix (the global-memory kernel), or for
//       classes up to kBatchResidentN 0: the resident kernel (four
//       wavefronts, the matrix in LDS) or kBatchTeamSym: the
//       symmetric-workspace kernel.  A batch runs with the mode of its
//       largest matrix's class.
//   ts  0: SIMT 4 x 4 register tiles; 16 or 32: MFMA tiles per wavefront
//       (ts x ts outputs); the resident and symmetric-workspace kernels use
//       MFMA (0 counts as 16).
struct BatchWidths {
  int w[kBatchClasses];
  int nt[kBatchClasses];
  int ts[kBatchClasses];
};

// Default table (used until bbk_batched_tune or bbk_batched_set_*): what
// the autotuner chose on MI355X up to n = 256 (one-wavefront teams below
// 129; batches whose largest order is above 128 in the symmetric-workspace
// mode; MFMA 32 x 32 update tiles); the global-memory kernel with SIMT
// updates above 256.
constexpr BatchWidths kDefaultBatchWidthsF64 = {{16, 8, 16, 16, 32, 32, 32, 32},
                                                {64, 64, 64, kBatchTeamSym, 256, 256, 256, 256},
                                                {32, 32, 32, 32, 0, 0, 0, 0}};
constexpr BatchWidths kDefaultBatchWidthsF32 = {{8, 8, 16, 24, 32, 32, 32, 32},
                                                {64, 64, 64, kBatchTeamSym, 256, 256, 256, 256},
                                                {32, 32, 32, 32, 0, 0, 0, 0}};

// The schedule of the last prepared batch: queue order, sizes, counter.
struct BatchPlan {
  int count = -1;
  int elem = 0;
  std::vector<int> sizes;
  BatchWidths widths{};
  void* buf = nullptr;  // ints: order[count], n[count], queue counter
  size_t buf_bytes = 0;
  void* ws = nullptr;   // global panel rows for matrices beyond LDS
  size_t ws_bytes = 0;
  int64_t ws_stride = 0;
  int grid = 0;
  int nt = 64;          // threads per matrix team
  bool lockstep = false;
  bool resident = false;  // matrices held in LDS (largest n <= kBatchResidentN)
  bool sym = false;       // symmetric-workspace kernel (largest n <= kBatchResidentN)
  int lds_dyn = 0;        // dynamic LDS bytes per workgroup: sym kernel or lockstep panel
  int steps = 0;        // lockstep: panel steps (largest ceil(n / width))
  int64_t step_stride = 0;  // lockstep: ints per step in the tile prefix table
  int64_t ws_elems = 0;
  std::vector<int> step_tiles;  // lockstep: trailing tiles per step
};

// Free the device buffers of a plan copy owned by a graph.
void release_plan(BatchPlan* p);

}  // namespace bbk

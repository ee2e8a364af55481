// bbk_internal.hpp -- shared device helpers and launch plumbing for the
// gfx950 blocked Bunch-Kaufman kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bbk {

// Bunch-Kaufman growth constant (1 + sqrt(17)) / 8,
// src/block_ldlt_decomposition.cc:14.
constexpr double kAlpha = 0.6403882032022076;
constexpr double kDblEps = 2.220446049250313080847e-16;

// Panel kernel geometry: one thread owns one active column; a workgroup
// of four wavefronts holds the panel's W rows for its 256 columns in LDS.
constexpr int kPanelThreads = 256;
constexpr int kPanelMaxWidth = 128;  // largest panel width of any type
constexpr int kMaxPanelGroups = 1024;
constexpr int kMaxPanelGroupsLds = 256;  // panel grid limit (n <= 65536)

// 16-byte exchange record {tag, aux, value} of the panel kernel; a record
// is valid for exchange e when tag == e.
struct __attribute__((aligned(16))) XRec {
  int tag;
  int aux;
  double v;
};

// Widest panel whose W rows (width + 1) x 256 fit the 160 KB LDS.
template <typename T>
struct PanelCfg;
template <>
struct PanelCfg<double> {
  static constexpr int kMaxWidth = 64;   // 65 x 256 x 8 B = 130 KB
};
template <>
struct PanelCfg<float> {
  static constexpr int kMaxWidth = 128;  // 129 x 256 x 4 B = 129 KB
};

// Exchange area layout (doubles) used by the panel kernel's workgroups
// between grid barriers.
struct PanelExchange {
  static constexpr int kOmega = 0;                          // [G] partial omega
  static constexpr int kSigma = kOmega + kMaxPanelGroups;   // [G] partial sigma
  static constexpr int kScalars = kSigma + kMaxPanelGroups; // 16 scalars
  static constexpr int kColK = kScalars + 16;               // [b+2]
  static constexpr int kColK1 = kColK + kPanelMaxWidth + 2; // [b+2]
  static constexpr int kColR = kColK1 + kPanelMaxWidth + 2; // [b+2]
  static constexpr int kDoubles = kColR + kPanelMaxWidth + 2;
  // scalar slots
  static constexpr int sWkk = 0, sWk1 = 1, sWkr = 2, sCrr = 3, sCk = 4,
                       sCk1 = 5, sAkkOld = 6, sAk1Old = 7;
};

// Per-panel device state (ints).
struct PanelState {
  static constexpr int kNextStart = 0;  // in: this panel's start; out: next
  static constexpr int kWidth = 1;      // out: columns factored
  static constexpr int kNumSwaps = 2;   // out: interchanges recorded
  static constexpr int kStart = 3;      // out: this panel's start
  static constexpr int kInts = 4;
};

// Grid-wide barrier for a cooperatively launched grid: monotonic counter,
// target = G * (barrier index + 1).  Agent-scope release/acquire via
// __threadfence around one atomic per workgroup.
__device__ __forceinline__ void grid_barrier(unsigned* counter,
                                             unsigned target) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    while (__hip_atomic_load(counter, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
    }
    __threadfence();
  }
  __syncthreads();
}

// Index of the first maximum: larger value wins, ties go to the lower
// index (matches the strict '>' scan of src/block_ldlt_decomposition.cc:173).
__device__ __forceinline__ void argmax_combine(double& v, int& i, double v2,
                                               int i2) {
  if (v2 > v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}

__device__ __forceinline__ double wave_max_nan_ignoring(double v) {
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(v, off, 64);
    v = (v < o) ? o : v;
  }
  return v;
}

__device__ __forceinline__ void wave_argmax(double& v, int& i) {
  for (int off = 32; off > 0; off >>= 1) {
    double ov = __shfl_xor(v, off, 64);
    int oi = __shfl_xor(i, off, 64);
    argmax_combine(v, i, ov, oi);
  }
}

// Workgroup-wide reductions; every thread receives the result.  sv/si
// are LDS scratch of at least blockDim.x / 64 entries.
__device__ __forceinline__ void block_argmax(double& v, int& i, double* sv,
                                             int* si) {
  wave_argmax(v, i);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sv[w] = v;
    si[w] = i;
  }
  __syncthreads();
  v = sv[0];
  i = si[0];
  for (int t = 1; t < nw; ++t) argmax_combine(v, i, sv[t], si[t]);
  __syncthreads();
}

__device__ __forceinline__ double block_max_nan_ignoring(double v,
                                                         double* sv) {
  v = wave_max_nan_ignoring(v);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) sv[w] = v;
  __syncthreads();
  v = sv[0];
  for (int t = 1; t < nw; ++t) v = (v < sv[t]) ? sv[t] : v;
  __syncthreads();
  return v;
}

}  // namespace bbk

// bbk_internal.hpp -- shared device helpers and launch plumbing for the
// gfx950 blocked Bunch-Kaufman kernels.
#pragma once

#include <climits>

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bbk {

// Bunch-Kaufman growth constant (1 + sqrt(17)) / 8,
// src/block_ldlt_decomposition.cc:14.
constexpr double kAlpha = 0.6403882032022076;
constexpr double kDblEps = 2.220446049250313080847e-16;

// Panel kernel geometry: one thread owns one active column; a workgroup
// of four wavefronts holds the panel's W rows for its 256 columns in LDS.
// Orders whose grid of 256-column groups would not fit co-resident
// (n > 65536) use workgroups of eight wavefronts (512 columns each).
constexpr int kPanelThreads = 256;
constexpr int kPanelThreadsWide = 512;
constexpr int kPanelMaxWidth = 128;  // largest panel width of any type
constexpr int kMaxPanelGroups = 1024;
// panel grid limit: n <= 65536 with 256-column groups, n <= 131072 with
// 512-column groups
constexpr int kMaxPanelGroupsLds = 256;

// 16-byte exchange record {tag, aux, value} of the panel kernel; a record
// is valid for exchange e when tag == e.
struct __attribute__((aligned(16))) XRec {
  int tag;
  int aux;
  double v;
};

// Panel width per type and workgroup size.  The panel's W rows live in
// LDS, (kLdsRows + 1) x NT elements; once the panel passes row kRegRows,
// rows [0, kRegRows) move to registers (one column per thread) and the LDS
// rows are reused for the rest.  The register rows' products need no LDS
// traffic, so half of the panel's rows sit there.  The rows [0, kRegRows]
// sit in LDS before they move, so kLdsRows >= kRegRows.
template <typename T, int NT = kPanelThreads>
struct PanelCfg;
template <>
struct PanelCfg<double, 256> {
  static constexpr int kLdsRows = 64;  // 65 x 256 x 8 B = 130 KB
  static constexpr int kRegRows = 64;  // 128 VGPRs
  static constexpr int kMaxWidth = kLdsRows + kRegRows;
};
template <>
struct PanelCfg<float, 256> {
  static constexpr int kLdsRows = 64;  // 65 x 256 x 4 B = 65 KB
  static constexpr int kRegRows = 64;  // 64 VGPRs
  static constexpr int kMaxWidth = kLdsRows + kRegRows;
};
template <>
struct PanelCfg<double, 512> {
  static constexpr int kLdsRows = 32;  // 33 x 512 x 8 B = 132 KB
  static constexpr int kRegRows = 32;  // 64 VGPRs
  static constexpr int kMaxWidth = kLdsRows + kRegRows;
};
template <>
struct PanelCfg<float, 512> {
  static constexpr int kLdsRows = 64;  // 65 x 512 x 4 B = 130 KB
  static constexpr int kRegRows = 64;  // 64 VGPRs
  static constexpr int kMaxWidth = kLdsRows + kRegRows;
};

// Largest panel width of the variant with nt threads per workgroup.
template <typename T>
constexpr int panel_max_width(int nt) {
  return nt == kPanelThreadsWide ? PanelCfg<T, kPanelThreadsWide>::kMaxWidth
                                 : PanelCfg<T, kPanelThreads>::kMaxWidth;
}

// Per-panel device state (ints).
struct PanelState {
  static constexpr int kNextStart = 0;  // in: this panel's start; out: next
  static constexpr int kWidth = 1;      // out: columns factored
  static constexpr int kNumSwaps = 2;   // out: interchanges recorded
  static constexpr int kStart = 3;      // out: this panel's start
  static constexpr int kInts = 4;
};

typedef int v4i __attribute__((ext_vector_type(4)));

// Device-coherent (sc1) 16-byte record store: one vector store carries the
// tag and the value together, so a reader that sees the tag sees the value.
__device__ __forceinline__ void xput(XRec* p, int tag, int aux, double v) {
  const long long b = __double_as_longlong(v);
  v4i d = {tag, aux, static_cast<int>(b & 0xffffffffll), static_cast<int>(b >> 32)};
  asm volatile("global_store_dwordx4 %0, %1, off sc1" ::"v"(p), "v"(d) : "memory");
}

// Coherent 16-byte record read; the wait sits in the same asm block so the
// destination registers are never observed before the data arrives.
// Returns false (outputs untouched) unless the record carries `tag`.
__device__ __forceinline__ bool xget(const XRec* p, int tag, int* aux, double* v) {
  v4i d;
  asm volatile("global_load_dwordx4 %0, %1, off sc1\n\ts_waitcnt vmcnt(0)"
               : "=v"(d)
               : "v"(p)
               : "memory");
  if (d.x != tag) return false;
  *aux = d.y;
  *v = __longlong_as_double((static_cast<long long>(d.w) << 32) |
                            static_cast<unsigned int>(d.z));
  return true;
}

// Device status words of a handle (4 x 32 bits in device memory):
//   kStatusError  bit 1: a panel exchange wait ran past its time limit;
//                 bit 2: a triangular-sweep wait did (results invalid).
//                 Kept until the host reads and clears it.
//   kStatusInfo   first singular pivot of the last factorization, as
//                 LAPACK's info: column k + 1 of a 1x1 pivot with
//                 |d| <= DBL_EPSILON or a 2x2 block with det = 0
//                 (0xffffffff = none).
//   kStatusNonFinite  1 when the last factorization's D holds a NaN or
//                 an infinity.
constexpr int kStatusError = 0, kStatusInfo = 1, kStatusNonFinite = 2, kStatusWords = 4;
constexpr unsigned kErrPanelWait = 1u, kErrSweepWait = 2u;
constexpr unsigned kInfoNone = 0xffffffffu;
// Time limit of a launch's waits when the handle sets none (far beyond
// any wait between co-resident workgroups of one launch).
constexpr int kDefaultWaitMs = 10000;

// Bounded waits of a cooperative launch.  Every writer is a co-resident
// workgroup of the same launch, so a wait ends unless a record is lost; a
// deadline on the device's constant-rate clock turns such a loss into an
// error the host reports instead of a hang.  Past the deadline every wait
// of the launch returns after its first poll (missing values: aux INT_MAX,
// value 0, never a valid index), so the launch still terminates.  forced:
// every wait fails at once (host-forced, to test the error path).  All
// fields are launch-uniform; the waits are wavefront-uniform loops.
struct WaitGuard {
  unsigned* status;
  unsigned long long deadline;
  bool forced;
  __device__ __forceinline__ WaitGuard(unsigned* st, long long ticks)
      : status(st), deadline(wall_clock64() + static_cast<unsigned long long>(ticks < 0 ? 0 : ticks)),
        forced(ticks < 0) {}
  __device__ __forceinline__ void fail(unsigned bit) const {
    if (status) atomicOr(status + kStatusError, bit);
  }
};

// A record no exchange ever writes, whose tag (0) no exchange uses: a wait
// past its deadline re-targets its polls here, so the poll loop keeps a
// single exit (the common path costs what an unbounded loop costs).
__device__ const XRec kTimeoutRecord = {0, INT_MAX, 0.0};

// Poll one record until it carries `tag`, within the guard's deadline
// (every 64th poll reads the clock).  Past it, or at once when forced, the
// poll is re-targeted to kTimeoutRecord (aux INT_MAX, value 0) and the
// error is marked.
__device__ __forceinline__ double xwait(const XRec* p, int tag, int* aux, const WaitGuard& g) {
  double v;
  if (g.forced) {
    p = &kTimeoutRecord;
    tag = 0;
  }
  unsigned polls = 0;
  while (!xget(p, tag, aux, &v)) {
    __builtin_amdgcn_s_sleep(1);
    if ((++polls & 63u) == 0 && wall_clock64() > g.deadline) {
      p = &kTimeoutRecord;
      tag = 0;
    }
  }
  if (tag == 0) g.fail(kErrPanelWait);
  return v;
}

// An interchange sequence (cols[2q], cols[2q+1]), q < ns, held in LDS,
// composed into one gather over its ne = 2 ns entries: afterwards position
// cols[t] takes the value found before the sequence at position
// content[slot[t]] (slot[t] is the first entry naming the same position).
// Called by every thread of the workgroup.
__device__ __forceinline__ void compose_swaps(const int* cols, int ns, int* slot,
                                              int* content) {
  const int ne = 2 * ns;
  for (int t = threadIdx.x; t < ne; t += blockDim.x) {
    const int cval = cols[t];
    int f = t;
    for (int u = 0; u < t; ++u)
      if (cols[u] == cval) {
        f = u;
        break;
      }
    slot[t] = f;
    content[t] = cval;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < ns; ++q) {
      const int sa = slot[2 * q], sb = slot[2 * q + 1];
      const int tmp = content[sa];
      content[sa] = content[sb];
      content[sb] = tmp;
    }
  }
  __syncthreads();
}

// Wait until this wave's outstanding vector memory operations completed.
__device__ __forceinline__ void wait_vm() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Pivot arithmetic kept free of FMA contraction so decisions and the 2x2
// inverse round like the reference.
__device__ __forceinline__ int select_pivot_nc(double akk, double omega, double sigma,
                                               double arr) {
#pragma clang fp contract(off)
  if (akk * sigma - kAlpha * omega * omega >= 0.0) return 3;
  if (arr - kAlpha * sigma >= 0.0) return -1;
  return 2;
}

__device__ __forceinline__ void inverse_2x2_nc(double a, double b, double c, double* itl,
                                               double* ibr, double* ioff) {
#pragma clang fp contract(off)
  const double det = a * c - b * b;
  *itl = c / det;
  *ibr = a / det;
  *ioff = -b / det;
}

__device__ __forceinline__ void apply_2x2_nc(double itl, double ibr, double ioff, double w0,
                                             double w1, double* u0, double* u1) {
#pragma clang fp contract(off)
  *u0 = itl * w0 + ioff * w1;
  *u1 = ioff * w0 + ibr * w1;
}

__device__ __forceinline__ double safe_den(double d) {
  return fabs(d) <= kDblEps ? 0.0 : d;  // SafeDenominator
}

__device__ __forceinline__ double div_or_zero(double w, double den) {
  return den == 0.0 ? 0.0 : w / den;
}

// Index of the first maximum: larger value wins, ties go to the lower
// index (matches the strict '>' scan of src/block_ldlt_decomposition.cc:173).
__device__ __forceinline__ void argmax_combine(double& v, int& i, double v2,
                                               int i2) {
  const bool take = (v2 > v) | ((v2 == v) & (i2 < i));  // no short circuit
  v = take ? v2 : v;
  i = take ? i2 : i;
}

// Wavefront reductions by DPP lane exchanges (VALU operand modifiers, no
// LDS round trip): two quad permutes, the half-row and row mirrors, then the
// row broadcasts of lanes 15 and 31 leave the result in lane 63, which is
// read back uniformly.  A lane whose DPP source is not written keeps its own
// value (update_dpp's `old`), which every combine below leaves unchanged.
constexpr int kDppQuadXor1 = 0xB1;       // quad_perm [1, 0, 3, 2]
constexpr int kDppQuadXor2 = 0x4E;       // quad_perm [2, 3, 0, 1]
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppRowMirror = 0x140;
constexpr int kDppRowBcast15 = 0x142;    // with row mask 0xA
constexpr int kDppRowBcast31 = 0x143;    // with row mask 0xC

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_i32(int x) {
  return __builtin_amdgcn_update_dpp(x, x, CTRL, ROW_MASK, 0xf, false);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_i32<CTRL, ROW_MASK>(static_cast<int>(b & 0xffffffffll));
  const int hi = dpp_i32<CTRL, ROW_MASK>(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}
__device__ __forceinline__ double readlane_f64(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

// max with NaN never taken over a number already held (as (v < o) ? o : v)
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void max_step(double& v) {
  const double o = dpp_f64<CTRL, ROW_MASK>(v);
  v = (v < o) ? o : v;
}
__device__ __forceinline__ double wave_max_nan_ignoring(double v) {
  max_step<kDppQuadXor1, 0xf>(v);
  max_step<kDppQuadXor2, 0xf>(v);
  max_step<kDppRowHalfMirror, 0xf>(v);
  max_step<kDppRowMirror, 0xf>(v);
  max_step<kDppRowBcast15, 0xa>(v);
  max_step<kDppRowBcast31, 0xc>(v);
  return readlane_f64(v, 63);
}

// first maximum: larger value wins, ties go to the lower index; branch-free
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void argmax_step(double& v, int& i) {
  const double ov = dpp_f64<CTRL, ROW_MASK>(v);
  const int oi = dpp_i32<CTRL, ROW_MASK>(i);
  const bool take = (ov > v) | ((ov == v) & (oi < i));  // no short circuit
  v = take ? ov : v;
  i = take ? oi : i;
}
__device__ __forceinline__ void wave_argmax(double& v, int& i) {
  argmax_step<kDppQuadXor1, 0xf>(v, i);
  argmax_step<kDppQuadXor2, 0xf>(v, i);
  argmax_step<kDppRowHalfMirror, 0xf>(v, i);
  argmax_step<kDppRowMirror, 0xf>(v, i);
  argmax_step<kDppRowBcast15, 0xa>(v, i);
  argmax_step<kDppRowBcast31, 0xc>(v, i);
  v = readlane_f64(v, 63);
  i = __builtin_amdgcn_readlane(i, 63);
}

// Forms for values that are never NaN (the callers map NaN to 0 with index
// INT_MAX): the maximum by DPP steps of v_max_f64, then the first index of
// it as the lowest index among the lanes holding it (one lane in the common
// case, read directly).  Same result as wave_argmax on such values.
__device__ __forceinline__ double wave_max_nonan(double v) {
  v = fmax(v, dpp_f64<kDppQuadXor1, 0xf>(v));
  v = fmax(v, dpp_f64<kDppQuadXor2, 0xf>(v));
  v = fmax(v, dpp_f64<kDppRowHalfMirror, 0xf>(v));
  v = fmax(v, dpp_f64<kDppRowMirror, 0xf>(v));
  v = fmax(v, dpp_f64<kDppRowBcast15, 0xa>(v));
  v = fmax(v, dpp_f64<kDppRowBcast31, 0xc>(v));
  return readlane_f64(v, 63);
}
__device__ __forceinline__ void wave_argmax_nonan(double& v, int& i) {
  const double m = wave_max_nonan(v);
  const bool hold = v == m;
  const unsigned long long hit = __builtin_amdgcn_ballot_w64(hold);
  int im;
  if (__builtin_popcountll(hit) == 1) {
    im = __builtin_amdgcn_readlane(i, __builtin_ctzll(hit));
  } else {
    int c = hold ? i : INT_MAX;
    c = min(c, dpp_i32<kDppQuadXor1, 0xf>(c));
    c = min(c, dpp_i32<kDppQuadXor2, 0xf>(c));
    c = min(c, dpp_i32<kDppRowHalfMirror, 0xf>(c));
    c = min(c, dpp_i32<kDppRowMirror, 0xf>(c));
    c = min(c, dpp_i32<kDppRowBcast15, 0xa>(c));
    c = min(c, dpp_i32<kDppRowBcast31, 0xc>(c));
    im = __builtin_amdgcn_readlane(c, 63);
  }
  v = m;
  i = im;
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

// One-barrier forms: the scratch alternates between two halves (sv, si
// hold 2 * blockDim.x / 64 entries; `parity` flips on every call), so a
// later reduction never overwrites values an earlier one may still be
// reading -- the barrier of the reduction in between orders them.
// (block_argmax1, block_max1: values never NaN)
__device__ __forceinline__ void block_argmax1(double& v, int& i, double* sv, int* si,
                                              int& parity) {
  wave_argmax_nonan(v, i);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double* svp = sv + parity * nw;
  int* sip = si + parity * nw;
  if ((threadIdx.x & 63) == 0) {
    svp[w] = v;
    sip[w] = i;
  }
  __syncthreads();
  v = svp[0];
  i = sip[0];
  for (int t = 1; t < nw; ++t) argmax_combine(v, i, svp[t], sip[t]);
  parity ^= 1;
}

__device__ __forceinline__ double block_max1(double v, double* sv, int& parity) {
  v = wave_max_nonan(v);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double* svp = sv + parity * nw;
  if ((threadIdx.x & 63) == 0) svp[w] = v;
  __syncthreads();
  v = svp[0];
  for (int t = 1; t < nw; ++t) v = (v < svp[t]) ? svp[t] : v;
  parity ^= 1;
  return v;
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

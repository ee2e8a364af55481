// bbk_batched.hip -- many independent Bunch-Kaufman LDL^T factorizations
// in one persistent launch.
//
// Each workgroup -- one wavefront when every matrix of the batch has
// n <= 256, so that more matrices are in flight, else four -- pulls matrix
// numbers from a device work queue (one atomic counter; the host orders the
// queue largest first) and factors the whole matrix by itself with the
// reference's blocked loop
// (src/block_ldlt_decomposition.cc:50-59,80-120): per panel, each pivot
// row is formed from the panel's unscaled rows W = D L^T, the
// Bunch-Kaufman tests pick a 1x1, a swapped 1x1 or a 2x2 pivot
// (:369-390), the symmetric interchanges touch the factored rows and the
// panel's W rows (:414-587), the pivot rows are scaled (:633-704), and a
// 2x2 that would end on the panel's last column is deferred to the next
// panel, which then starts one column earlier (:589-631).  After the
// panel, the trailing upper triangle is updated, C -= U^T W (:706-744).
// All synchronization is inside the workgroup (__syncthreads), so matrices
// proceed independently.
//
// Panel width comes per matrix from a size-class table (tuned by
// bbk_batched_tune).  The panel rows W (width + 1 rows) and the candidate
// row live in LDS when they fit the dynamic LDS budget, else in a global
// slice owned by the workgroup; both are reached through one generic
// pointer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>

#include "../../include/bbk_ldlt.h"
#include "bbk_batched.hpp"
#include "bbk_handle.hpp"
#include "bbk_update.hpp"

namespace bbk {

template <typename T>
struct BatchArgs {
  int count;
  const int* order;     // queue position -> matrix number
  const int* ns;        // matrix number -> order n
  T* const* A;
  const int64_t* lda;
  int* const* perm;
  int* const* piv;
  int* queue;           // work counter, zero at launch
  T* ws;                // global W slices (gridDim.x of ws_stride elements)
  int64_t ws_stride;
  int lds_elems;        // dynamic LDS capacity in elements of T
  int lds_bytes;        // resident kernel: dynamic LDS arena bytes
  BatchWidths widths;
};

extern __shared__ __attribute__((aligned(16))) unsigned char bbk_batched_lds[];

__host__ __device__ __forceinline__ int batch_class(int n) {
  int c = 0;
  while (c < kBatchClasses - 1 && n > batch_class_bound(c)) ++c;
  return c;
}

template <typename T>
__device__ __forceinline__ void swap_elems(T* a, T* b) {
  const T t = *a;
  *a = *b;
  *b = t;
}

// Reductions over the team (one wavefront or the whole workgroup); every
// thread receives the result.
template <int NT>
__device__ __forceinline__ void team_argmax(double& v, int& i, double* sv, int* si) {
  if constexpr (NT == 64) {
    wave_argmax(v, i);
  } else {
    block_argmax(v, i, sv, si);
  }
}
template <int NT>
__device__ __forceinline__ double team_max(double v, double* sv) {
  if constexpr (NT == 64) {
    return wave_max_nan_ignoring(v);
  } else {
    return block_max_nan_ignoring(v, sv);
  }
}

// Per-panel scaling record of row p (LDS): ty 1: U = W / c0 (c0 =
// SafeDenominator(d)); ty 2: U = c0 W(p) + c1 W(p+1); ty -2: U = c0 W(p-1) +
// c1 W(p).  U(p, col) rebuilt from W equals the value the scaling step wrote
// to F(s + p, col) bit for bit (same operands, same operations), also after
// later column interchanges, which move F and W columns together.
struct ScaleTab {
  int ty[kBatchMaxWidth + 2];
  double c0[kBatchMaxWidth + 2];
  double c1[kBatchMaxWidth + 2];
};

template <typename T>
__device__ __forceinline__ T ucol(const ScaleTab& st, const T* W, int n, int p, int col) {
#pragma clang fp contract(off)
  const int ty = st.ty[p];
  const double w = static_cast<double>(W[(int64_t)p * n + col]);
  if (ty == 1) return static_cast<T>(div_or_zero(w, st.c0[p]));
  if (ty == 2)
    return static_cast<T>(st.c0[p] * w +
                          st.c1[p] * static_cast<double>(W[(int64_t)(p + 1) * n + col]));
  return static_cast<T>(st.c0[p] * static_cast<double>(W[(int64_t)(p - 1) * n + col]) +
                        st.c1[p] * w);
}

// sum_{p < kk} l[p] W(p, j) (row stride n).  ACC4: four interleaved
// partial sums, so that four loads of W are in flight per step (W in global
// memory, and binary64 in LDS: measured faster); else one running sum
// (binary32 with W in LDS: measured faster).
template <typename T, bool ACC4>
__device__ __forceinline__ T panel_rows_dot(const T* l, const T* W, int n, int kk, int j) {
  if constexpr (!ACC4) {
    T acc = T(0);
    for (int p = 0; p < kk; ++p) acc += l[p] * W[(int64_t)p * n + j];
    return acc;
  }
  T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
  int p = 0;
  for (; p + 4 <= kk; p += 4) {
    a0 += l[p] * W[(int64_t)p * n + j];
    a1 += l[p + 1] * W[(int64_t)(p + 1) * n + j];
    a2 += l[p + 2] * W[(int64_t)(p + 2) * n + j];
    a3 += l[p + 3] * W[(int64_t)(p + 3) * n + j];
  }
  for (; p < kk; ++p) a0 += l[p] * W[(int64_t)p * n + j];
  return (a0 + a1) + (a2 + a3);
}

// The products of PB positions j0, j0 + NT, ..., j0 + (PB - 1) NT at once
// (positions >= n give unused values): for each position the same
// operations in the same order as panel_rows_dot, so the results are
// bitwise those of PB separate calls, with PB times the loads of W in
// flight (W in global memory is latency-bound one position at a time).
template <typename T, bool ACC4, int PB, int NT>
__device__ __forceinline__ void panel_rows_dot_multi(const T* l, const T* W, int n, int kk,
                                                     int j0, T (&out)[PB],
                                                     const T* Wl = nullptr, int RL = 0) {
  int jj[PB];
#pragma unroll
  for (int u = 0; u < PB; ++u) jj[u] = (j0 + u * NT < n) ? j0 + u * NT : j0;  // j0 < n
  if constexpr (!ACC4) {
    T acc[PB];
#pragma unroll
    for (int u = 0; u < PB; ++u) acc[u] = T(0);
    for (int p = 0; p < kk; ++p) {
      const T lp = l[p];
      const T* row = W + (int64_t)p * n;
#pragma unroll
      for (int u = 0; u < PB; ++u) acc[u] += lp * row[jj[u]];
    }
#pragma unroll
    for (int u = 0; u < PB; ++u) out[u] = acc[u];
    return;
  }
  T a0[PB], a1[PB], a2[PB], a3[PB];
#pragma unroll
  for (int u = 0; u < PB; ++u) a0[u] = a1[u] = a2[u] = a3[u] = T(0);
  int p = 0;
  // rows [0, RL) (RL a multiple of 4) from the copy in LDS: the same
  // values, summed in the same order
  const int kl = min(kk, RL);
  for (; p + 4 <= kl; p += 4) {
    const T* r0 = Wl + p * n;
    const T* r1 = r0 + n;
    const T* r2 = r1 + n;
    const T* r3 = r2 + n;
#pragma unroll
    for (int u = 0; u < PB; ++u) {
      a0[u] += l[p] * r0[jj[u]];
      a1[u] += l[p + 1] * r1[jj[u]];
      a2[u] += l[p + 2] * r2[jj[u]];
      a3[u] += l[p + 3] * r3[jj[u]];
    }
  }
  for (; p + 4 <= kk; p += 4) {
    const T* r0 = W + (int64_t)p * n;
    const T* r1 = r0 + n;
    const T* r2 = r1 + n;
    const T* r3 = r2 + n;
#pragma unroll
    for (int u = 0; u < PB; ++u) {
      a0[u] += l[p] * r0[jj[u]];
      a1[u] += l[p + 1] * r1[jj[u]];
      a2[u] += l[p + 2] * r2[jj[u]];
      a3[u] += l[p + 3] * r3[jj[u]];
    }
  }
  for (; p < kk; ++p) {
    const T* row = W + (int64_t)p * n;
#pragma unroll
    for (int u = 0; u < PB; ++u) a0[u] += l[p] * row[jj[u]];
  }
#pragma unroll
  for (int u = 0; u < PB; ++u) out[u] = (a0[u] + a1[u]) + (a2[u] + a3[u]);
}

// Storage views of one matrix for the batched factorization; element
// (i, j) with i <= j (upper triangle).
//   GlobalMat  row-major in global memory (leading dimension lda).
//   LdsMat     packed in LDS: row i holds columns i..n-1 from offset
//              i n - i (i - 1) / 2.
//   HybridMat  rows >= R0 packed in LDS as above (the trailing square
//              [R0, n)^2 on chip), rows < R0 in global memory.  In a panel
//              that starts at or after R0, the interchanges that reach the
//              rows < R0 are recorded (defer) and applied after the last
//              panel: those rows belong to finished panels, which nothing
//              reads again (L rows come from W, the trailing update reads
//              only the current panel's rows).  first_immediate_row(s): the
//              first row an interchange in the panel starting at s updates
//              at once.
template <typename T>
struct GlobalMat {
  T* g;
  int64_t lda;
  __device__ __forceinline__ T& at(int i, int j) const { return g[(int64_t)i * lda + j]; }
  __device__ __forceinline__ int first_immediate_row(int) const { return 0; }
  __device__ __forceinline__ void defer(int, int) const {}
};
// LDS record of one panel's deferred interchanges (at most one per column
// of a panel of up to kBatchMaxWidth + 1 columns) and the scratch of their
// composition.
struct DeferScratch {
  static constexpr int kMax = kBatchMaxWidth + 2;
  int list[2 * kMax];
  int slot[2 * kMax];
  int content[2 * kMax];
  int src[2 * kMax];
  int count;
};

// GlobalMat whose interchanges reach only the rows at or after the panel
// start at once; those of the finished rows above it are recorded (defer,
// into a DeferScratch list) and applied after the panel as one composed
// gather per row (apply_panel_deferred), which touches each row once
// instead of once per interchange.
template <typename T>
struct DeferGlobalMat {
  T* g;
  int64_t lda;
  DeferScratch* ds;
  __device__ __forceinline__ T& at(int i, int j) const { return g[(int64_t)i * lda + j]; }
  __device__ __forceinline__ int first_immediate_row(int s) const { return s; }
  // called by one thread
  __device__ __forceinline__ void defer(int a, int b) const {
    const int c = ds->count;
    ds->list[2 * c] = a;
    ds->list[2 * c + 1] = b;
    ds->count = c + 1;
  }
};
template <typename T>
struct LdsMat {
  T* s;
  int n;
  __device__ __forceinline__ T& at(int i, int j) const {
    return s[i * n - ((i * (i - 1)) >> 1) + (j - i)];
  }
  __device__ __forceinline__ int first_immediate_row(int) const { return 0; }
  __device__ __forceinline__ void defer(int, int) const {}
};
template <typename T>
struct HybridMat {
  T* g;
  int64_t lda;
  T* s;
  int n, R0;
  int* dlist;   // deferred interchanges (a, b), in order
  int* dcount;
  __device__ __forceinline__ T& at(int i, int j) const {
    if (i < R0) return g[(int64_t)i * lda + j];
    const int r = i - R0, m = n - R0;
    return s[r * m - ((r * (r - 1)) >> 1) + (j - i)];
  }
  __device__ __forceinline__ int first_immediate_row(int s) const { return s >= R0 ? R0 : 0; }
  // called by one thread
  __device__ __forceinline__ void defer(int a, int b) const {
    const int c = *dcount;
    dlist[2 * c] = a;
    dlist[2 * c + 1] = b;
    *dcount = c + 1;
  }
};

// Panel P of one matrix, whole team of NT threads, starting at column s
// (one column early when the previous panel deferred a 2x2): Bunch-Kaufman
// steps with the panel rows W (row stride n; rows 0..bw = panel rows, row
// bw + 1 = the candidate row C).  F: the matrix's storage view (GlobalMat,
// LdsMat or HybridMat); perm/piv in global memory or LDS.  Returns the
// panel's end e (the nominal end, or the column of a deferred 2x2).  PB:
// positions per thread whose row products are formed together
// (panel_rows_dot_multi; 1 = one at a time).  LROW: the current pivot's W
// row and the candidate row live in LDS (rowk, rowc: n elements each,
// indexed by column) instead of W's rows kk and bw + 1, and each finished
// row is copied to W after its scaling; for W in global memory this takes
// the row reads between the pivot's phases off global memory.  With LROW,
// Wl (LDS, RL rows of n, RL a multiple of 4; ACC4 only) also holds W's rows
// [0, RL) once finished, interchanges applied to both copies, and the row
// products read those rows from it.
template <typename T, int NT, bool ACC4, int PB = 1, bool LROW = false, typename Mat>
__device__ int factor_panel(const Mat& F, int n, int* __restrict__ perm,
                            int* __restrict__ piv, int bw, int P, int s, T* W, T* Lk, T* Lr,
                            ScaleTab& st, double* rv, int* ri, T* rowk = nullptr,
                            T* rowc = nullptr, T* Wl = nullptr, int RL = 0) {
  static_assert(ACC4 || !LROW, "the LDS copy of W's rows is read by the four-sum products");
  const int tid = threadIdx.x;
  T* C = LROW ? rowc : W + (int64_t)(bw + 1) * n;
  int e = min((P + 1) * bw, n);  // nominal end; s may sit one column early
  int kk = 0;
  while (s + kk < e) {
    const int k = s + kk;
    if (tid < kk) Lk[tid] = ucol<T>(st, W, n, tid, k);  // L(k, s + tid)
    __syncthreads();
    // W(k, j) = F(k, j) - sum_p L(k, p) W(p, j), j >= k  (:122-162)
    double best = 0.0;
    int arg = n;
    T* Wk = LROW ? rowk : W + (int64_t)kk * n;
    for (int j0 = k + tid; j0 < n; j0 += PB * NT) {
      T fk[PB], dot[PB];
#pragma unroll
      for (int u = 0; u < PB; ++u) fk[u] = (j0 + u * NT < n) ? F.at(k, j0 + u * NT) : T(0);
      panel_rows_dot_multi<T, ACC4, PB, NT>(Lk, W, n, kk, j0, dot, Wl, RL);
#pragma unroll
      for (int u = 0; u < PB; ++u) {
        const int j = j0 + u * NT;
        if (j < n) {
          const T w = fk[u] - dot[u];
          Wk[j] = w;
          const double a = fabs(static_cast<double>(w));
          if (j > k && a > best) {
            best = a;
            arg = j;
          }
        }
      }
    }
    team_argmax<NT>(best, arg, rv, ri);  // (:164-182) first maximum
    __syncthreads();
    const double omega = best;
    const int r = arg;
    const double akk = fabs(static_cast<double>(Wk[k]));
    int kind = 1;
    if (!(akk >= kAlpha * omega || omega == 0.0)) {
      // candidate row r: C(i) = A(r, i) - sum_p L(r, p) W(p, i), i >= k
      // (:184-352), sigma over i != r (:354-367)
      if (tid < kk) Lr[tid] = ucol<T>(st, W, n, tid, r);  // L(r, s + tid)
      __syncthreads();
      double sig = 0.0;
      for (int i0 = k + tid; i0 < n; i0 += PB * NT) {
        T base[PB], dot[PB];
#pragma unroll
        for (int u = 0; u < PB; ++u) {
          const int i = i0 + u * NT;
          base[u] = i < n ? (i < r ? F.at(i, r) : F.at(r, i)) : T(0);
        }
        panel_rows_dot_multi<T, ACC4, PB, NT>(Lr, W, n, kk, i0, dot, Wl, RL);
#pragma unroll
        for (int u = 0; u < PB; ++u) {
          const int i = i0 + u * NT;
          if (i < n) {
            const T cv = base[u] - dot[u];
            C[i] = cv;
            const double a = fabs(static_cast<double>(cv));
            if (i != r) sig = (sig < a) ? a : sig;
          }
        }
      }
      const double sigma = team_max<NT>(sig, rv);
      __syncthreads();
      const double arr = fabs(static_cast<double>(C[r]));
      kind = select_pivot_nc(akk, omega, sigma, arr);  // 3, -1 or 2
      if (kind == 2 && kk == e - s - 1 && e < n) {
        // 2x2 across the panel edge: this panel ends before k
        e = k;
        __syncthreads();
        break;
      }
      // the candidate's value at k replaces W(k, r) (column part of the
      // candidate row written over row k's entry)
      if (tid == 0) Wk[r] = C[k];
      __syncthreads();
      if (kind == -1) {
        // interchange k <-> r (:414-498); row k of W becomes the candidate
        for (int j = r + 1 + tid; j < n; j += NT) swap_elems(&F.at(k, j), &F.at(r, j));
        for (int j = k + 1 + tid; j < r; j += NT) swap_elems(&F.at(k, j), &F.at(j, r));
        // rows above k: rows the view holds off chip wait until the end
        const int i0 = F.first_immediate_row(s);
        for (int i = i0 + tid; i < k; i += NT) swap_elems(&F.at(i, k), &F.at(i, r));
        if (i0 > 0 && tid == 0) F.defer(k, r);
        for (int p = tid; p < kk; p += NT)
          swap_elems(&W[(int64_t)p * n + k], &W[(int64_t)p * n + r]);
        for (int p = tid; p < min(kk, RL); p += NT) swap_elems(&Wl[p * n + k], &Wl[p * n + r]);
        for (int j = k + 1 + tid; j < n; j += NT)
          if (j != r) Wk[j] = C[j];
        if (tid == 0) {
          swap_elems(&F.at(k, k), &F.at(r, r));
          Wk[k] = C[r];
          const int t = perm[k];
          perm[k] = perm[r];
          perm[r] = t;
        }
      } else if (kind == 2) {
        // interchange k+1 <-> r (:500-587); row k+1 of W becomes the candidate
        const int k1 = k + 1;
        // row k+1 of W: with LROW the candidate row itself (C)
        T* Wk1 = LROW ? C : W + (int64_t)(kk + 1) * n;
        if (r != k1) {
          for (int j = r + 1 + tid; j < n; j += NT) swap_elems(&F.at(k1, j), &F.at(r, j));
          for (int j = k + 2 + tid; j < r; j += NT) swap_elems(&F.at(k1, j), &F.at(j, r));
          const int i0 = F.first_immediate_row(s);
          for (int i = i0 + tid; i < k; i += NT) swap_elems(&F.at(i, k1), &F.at(i, r));
          if (i0 > 0 && tid == 0) F.defer(k1, r);
          for (int p = tid; p < kk; p += NT)
            swap_elems(&W[(int64_t)p * n + k1], &W[(int64_t)p * n + r]);
          for (int p = tid; p < min(kk, RL); p += NT)
            swap_elems(&Wl[p * n + k1], &Wl[p * n + r]);
        }
        if constexpr (!LROW) {
          for (int j = k + 2 + tid; j < n; j += NT)
            if (j != r) Wk1[j] = C[j];
        }
        if (tid == 0) {
          if (r != k1) {
            swap_elems(&F.at(k1, k1), &F.at(r, r));
            swap_elems(&F.at(k, k1), &F.at(k, r));
            swap_elems(&Wk[k1], &Wk[r]);
            // W(k+1, r) = C(k+1), W(k+1, k+1) = C(r): in place a swap
            if constexpr (LROW)
              swap_elems(&C[k1], &C[r]);
            else
              Wk1[r] = C[k1];
            const int t = perm[k1];
            perm[k1] = perm[r];
            perm[r] = t;
          }
          if constexpr (!LROW) Wk1[k1] = C[r];
        }
      }
      __syncthreads();
    }
    if (kind != 2) {
      // ScaleOneByOnePivot (:633-656)
      const double d = static_cast<double>(Wk[k]);
      const double den = safe_den(d);
      for (int j = k + 1 + tid; j < n; j += NT)
        F.at(k, j) = static_cast<T>(div_or_zero(static_cast<double>(Wk[j]), den));
      if (tid == 0) {
        F.at(k, k) = Wk[k];
        piv[k] = 1;
        st.ty[kk] = 1;
        st.c0[kk] = den;
      }
      if constexpr (LROW) {
        T* g = W + (int64_t)kk * n;
        for (int j = k + tid; j < n; j += NT) g[j] = Wk[j];
        if (kk < RL)
          for (int j = k + tid; j < n; j += NT) Wl[kk * n + j] = Wk[j];
      }
      kk += 1;
    } else {
      // ScaleTwoByTwoPivot (:658-704), binary64 without contraction
      const int k1 = k + 1;
      const T* Wk1 = LROW ? C : W + (int64_t)(kk + 1) * n;
      const double a = static_cast<double>(Wk[k]);
      const double b = static_cast<double>(Wk[k1]);
      const double c = static_cast<double>(Wk1[k1]);
      double itl, ibr, ioff;
      inverse_2x2_nc(a, b, c, &itl, &ibr, &ioff);
      for (int j = k + 2 + tid; j < n; j += NT) {
        double u0, u1;
        apply_2x2_nc(itl, ibr, ioff, static_cast<double>(Wk[j]), static_cast<double>(Wk1[j]),
                     &u0, &u1);
        F.at(k, j) = static_cast<T>(u0);
        F.at(k1, j) = static_cast<T>(u1);
      }
      if (tid == 0) {
        F.at(k, k) = Wk[k];
        F.at(k, k1) = Wk[k1];
        F.at(k1, k1) = Wk1[k1];
        piv[k] = 2;
        piv[k1] = -2;
        st.ty[kk] = 2;
        st.c0[kk] = itl;
        st.c1[kk] = ioff;
        st.ty[kk + 1] = -2;
        st.c0[kk + 1] = ioff;
        st.c1[kk + 1] = ibr;
      }
      if constexpr (LROW) {
        T* g0 = W + (int64_t)kk * n;
        T* g1 = g0 + n;
        for (int j = k + tid; j < n; j += NT) {
          g0[j] = Wk[j];
          g1[j] = Wk1[j];
        }
        if (kk < RL)
          for (int j = k + tid; j < n; j += NT) Wl[kk * n + j] = Wk[j];
        if (kk + 1 < RL)
          for (int j = k + tid; j < n; j += NT) Wl[(kk + 1) * n + j] = Wk1[j];
      }
      kk += 2;
    }
    __syncthreads();
  }
  return e;
}

// Trailing update F(i, j) -= sum_p F(s+p, i) W(p, j), e <= i <= j < n
// (:706-744, src/kernels.cc:67-91), 4 x 4 register tiles per thread.
template <typename T, int NT>
__device__ void simt_update(T* __restrict__ F, int64_t lda, int n, int s, int e, const T* W) {
  const int tid = threadIdx.x;
  const int m = n - e, kp = e - s;
  if (m > 0 && kp > 0) {
    const int nt = (m + 3) >> 2;
    const int tiles = nt * (nt + 1) / 2;
    for (int t = tid; t < tiles; t += NT) {
      int tj = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
      while (tj * (tj + 1) / 2 > t) --tj;
      while ((tj + 1) * (tj + 2) / 2 <= t) ++tj;
      const int ti = t - tj * (tj + 1) / 2;
      const int i0 = e + 4 * ti, j0 = e + 4 * tj;
      T acc[4][4];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = T(0);
      for (int p = 0; p < kp; ++p) {
        const T* urow = F + (int64_t)(s + p) * lda;
        const T* wrow = W + (int64_t)p * n;
        T u[4], w[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          u[a] = i0 + a < n ? urow[i0 + a] : T(0);
          w[a] = j0 + a < n ? wrow[j0 + a] : T(0);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[a][b] += u[a] * w[b];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int i = i0 + a, j = j0 + b;
          if (i <= j && j < n) F[(int64_t)i * lda + j] -= acc[a][b];
        }
    }
  }
}

// MFMA trailing update C(i, j) -= sum_{p < kp} U(p, i) W(p, j), e <= i <= j < n,
// U(p, i) = F(s + p, i), kp = e - s (src/block_ldlt_decomposition.cc:706-744,
// src/kernels.cc:67-91) on MFMA: each wavefront takes TS x TS output tiles
// of the upper triangle (TS = 16: one 16x16x4 block; 32: 2 x 2 blocks
// sharing their operands), K in steps of 4.  U is negated as it is loaded,
// so the accumulators started from C hold C - U^T W.
template <typename T, int NT, int TS, typename Mat>
__device__ void mfma_update(const Mat& F, int n, int s, int e, const T* W) {
  constexpr int TB = TS / 16;
  typedef typename Mfma<T>::acc_t acc_t;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = n - e, kp = e - s;
  if (m <= 0 || kp <= 0) return;
  const int nt = (m + TS - 1) / TS, tiles = nt * (nt + 1) / 2;
  for (int t = wave; t < tiles; t += NT / 64) {
    int I, J;
    tri_tile(t, &I, &J);
    const int i0 = e + I * TS, j0 = e + J * TS;
    acc_t acc[TB][TB];
#pragma unroll
    for (int ti = 0; ti < TB; ++ti)
#pragma unroll
      for (int tj = 0; tj < TB; ++tj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = i0 + 16 * ti + Mfma<T>::row(lane, v), j = j0 + 16 * tj + (lane & 15);
          acc[ti][tj][v] = (i <= j && j < n) ? F.at(i, j) : T(0);
        }
    const int q = lane >> 4, c = lane & 15;
    for (int k4 = 0; k4 < kp; k4 += 4) {
      const int p = k4 + q;
      T a[TB], b[TB];
#pragma unroll
      for (int tb = 0; tb < TB; ++tb) {
        const int ia = i0 + 16 * tb + c, jb = j0 + 16 * tb + c;
        a[tb] = (p < kp && ia < n) ? -F.at(s + p, ia) : T(0);
        b[tb] = (p < kp && jb < n) ? W[p * n + jb] : T(0);
      }
#pragma unroll
      for (int ti = 0; ti < TB; ++ti)
#pragma unroll
        for (int tj = 0; tj < TB; ++tj) acc[ti][tj] = Mfma<T>::mma(a[ti], b[tj], acc[ti][tj]);
    }
#pragma unroll
    for (int ti = 0; ti < TB; ++ti)
#pragma unroll
      for (int tj = 0; tj < TB; ++tj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = i0 + 16 * ti + Mfma<T>::row(lane, v), j = j0 + 16 * tj + (lane & 15);
          if (i <= j && j < n) F.at(i, j) = acc[ti][tj][v];
        }
  }
}

// One matrix, whole team: every panel, then its trailing update (ts: 0 =
// SIMT 4 x 4 register tiles, 16 or 32 = MFMA tiles per wavefront).
template <typename T, int NT>
__device__ void factor_one(T* __restrict__ F, int64_t lda, int n, int* __restrict__ perm,
                           int* __restrict__ piv, int bw, int ts, T* W, T* Lk, T* Lr,
                           ScaleTab& st, double* rv, int* ri) {
  for (int i = threadIdx.x; i < n; i += NT) {
    perm[i] = i;
    piv[i] = 1;
  }
  __syncthreads();
  int s = 0;
  for (int P = 0; s < n; ++P) {
    // persistent teams: the matrix's W in LDS unless it is large (binary64
    // past about 280): four partial sums for binary64 only
    const int e = factor_panel<T, NT, sizeof(T) == 8>(GlobalMat<T>{F, lda}, n, perm, piv, bw, P,
                                                       s, W, Lk, Lr, st, rv, ri);
    if (ts == 16)
      mfma_update<T, NT, 16>(GlobalMat<T>{F, lda}, n, s, e, W);
    else if (ts == 32)
      mfma_update<T, NT, 32>(GlobalMat<T>{F, lda}, n, s, e, W);
    else
      simt_update<T, NT>(F, lda, n, s, e, W);
    __syncthreads();
    s = e;
  }
}

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_batched_sytrf(BatchArgs<T> a) {
  __shared__ T Lk[kBatchMaxWidth + 2];
  __shared__ T Lr[kBatchMaxWidth + 2];
  __shared__ ScaleTab st;
  __shared__ double rv[NT / 64];
  __shared__ int ri[NT / 64];
  __shared__ int sq;
  T* lds = reinterpret_cast<T*>(bbk_batched_lds);
  for (;;) {
    if (threadIdx.x == 0) sq = atomicAdd(a.queue, 1);
    __syncthreads();
    const int q = sq;
    __syncthreads();
    if (q >= a.count) break;
    const int m = a.order[q];
    const int n = a.ns[m];
    if (n <= 0) continue;
    const int c = batch_class(n);
    const int bw = a.widths.w[c];
    const int64_t need = (int64_t)(bw + 2) * n;
    T* W = need <= a.lds_elems ? lds : a.ws + (int64_t)blockIdx.x * a.ws_stride;
    factor_one<T, NT>(a.A[m], a.lda[m], n, a.perm[m], a.piv[m], bw, a.widths.ts[c], W, Lk, Lr, st,
                      rv, ri);
  }
}

// Panel P of one matrix held whole in LDS (packed upper triangle S, order
// n <= 256 = NT: thread j owns column j for the whole factorization).  The
// same Bunch-Kaufman steps and arithmetic as factor_panel
// (src/block_ldlt_decomposition.cc:80-120,122-704), fused so that a pivot
// meets at most three barriers:
//   A  row k: W(k, j) = S(k, j) - sum_p L(k, p) W(p, j), the first argmax
//      of |W(k, j)|, j > k (one reduction);
//   B  (when the test needs it) candidate row r and its maximum sigma over
//      i != r (one reduction);
//   C  interchange and scaling in one pass: the interchange's writes into
//      the pivot row(s) are dead (the scaling overwrites them), so only the
//      moves into row/column r, the earlier rows' column swaps and the W
//      column swaps remain, each element touched by one thread.
// L(k, s + p) is read straight from the finished rows, S(s + p, k): the
// stored U values are the scaled W values the reference uses.  Values a
// thread needs from another column (W(k, k) with its sign, C(k), C(k+1),
// C(r), the old W(k, k+1) and W(k, r)) go through LDS slots written before
// the reduction barrier that precedes their reads.
struct PivotSlots {
  double wkk, ck, ck1, crr;
};

template <typename T>
__device__ __forceinline__ int pk(int n, int i, int j) {
  return i * n - ((i * (i - 1)) >> 1) + (j - i);
}

// sum_{p < kk} S(s + p, col) W(p, j), four interleaved partial sums for
// binary64 (as panel_rows_dot), one running sum for binary32.
template <typename T, bool ACC4>
__device__ __forceinline__ T lds_rows_dot(const T* S, int n, int s, int col, const T* W, int kk,
                                          int j) {
  int o = pk<T>(n, s, col);
  int step = n - s - 1;  // pk(i + 1, col) - pk(i, col) = n - i - 1
  if constexpr (!ACC4) {
    T acc = T(0);
    for (int p = 0; p < kk; ++p) {
      acc += S[o] * W[p * n + j];
      o += step;
      --step;
    }
    return acc;
  }
  T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
  int p = 0;
  for (; p + 4 <= kk; p += 4) {
    const int o1 = o + step, o2 = o1 + step - 1, o3 = o2 + step - 2;
    a0 += S[o] * W[p * n + j];
    a1 += S[o1] * W[(p + 1) * n + j];
    a2 += S[o2] * W[(p + 2) * n + j];
    a3 += S[o3] * W[(p + 3) * n + j];
    o = o3 + step - 3;
    step -= 4;
  }
  for (; p < kk; ++p) {
    a0 += S[o] * W[p * n + j];
    o += step;
    --step;
  }
  return (a0 + a1) + (a2 + a3);
}

template <typename T, bool ACC4>
__device__ int factor_panel_lds(T* S, int n, int* perm, int* piv, int bw, int P, int s, T* W,
                                PivotSlots& ps, double* rv, int* ri, int& par) {
  const int j = threadIdx.x;  // owned column
  T* C = W + (bw + 1) * n;
  int e = min((P + 1) * bw, n);
  int kk = 0;
  while (s + kk < e) {
    const int k = s + kk;
    T* Wk = W + kk * n;
    // ---- A: row k -------------------------------------------------------
    T w = T(0);
    double v = 0.0;
    int vi = INT_MAX;
    if (j >= k && j < n) {
      w = S[pk<T>(n, k, j)] - lds_rows_dot<T, ACC4>(S, n, s, k, W, kk, j);
      Wk[j] = w;
      const double a = fabs(static_cast<double>(w));
      if (j > k && a > 0.0) {  // NaN never wins (strict '>' scan, :164-182)
        v = a;
        vi = j;
      }
      if (j == k) ps.wkk = static_cast<double>(w);
    }
    block_argmax1(v, vi, rv, ri, par);
    const double omega = v;
    const int r = vi;
    const double wkk = ps.wkk;
    const double akk = fabs(wkk);
    int kind = 1;
    double ck = 0.0, ck1 = 0.0, crr = 0.0;
    T c = T(0);
    T wk_k1 = T(0);
    if (!(akk >= kAlpha * omega || omega == 0.0)) {
      // ---- B: candidate row r (:184-367) ---------------------------------
      if (k + 1 < n) wk_k1 = Wk[k + 1];
      double sv = 0.0;
      if (j >= k && j < n) {
        const T base = j < r ? S[pk<T>(n, j, r)] : S[pk<T>(n, r, j)];
        c = base - lds_rows_dot<T, ACC4>(S, n, s, r, W, kk, j);
        C[j] = c;
        const double a = fabs(static_cast<double>(c));
        if (j != r) sv = (a >= 0.0) ? a : 0.0;  // NaN ignored, as std::max does
        if (j == k) ps.ck = static_cast<double>(c);
        if (j == k + 1) ps.ck1 = static_cast<double>(c);
        if (j == r) ps.crr = static_cast<double>(c);
      }
      const double sigma = block_max1(sv, rv, par);
      crr = ps.crr;
      ck = ps.ck;
      ck1 = ps.ck1;
      kind = select_pivot_nc(akk, omega, sigma, fabs(crr));
      if (kind == 2 && kk == e - s - 1 && e < n) {
        e = k;  // 2x2 across the panel edge: deferred to the next panel (:589-631)
        break;
      }
    }
    // ---- C: interchange + scaling -----------------------------------------
    if (kind == 1 || kind == 3) {
      // ScaleOneByOnePivot (:633-656); after a candidate row that kept the
      // pivot, W(k, r) is the candidate's value at k (:184-247)
      const double den = safe_den(wkk);
      if (j > k && j < n) {
        T wj = w;
        if (kind == 3 && j == r) {
          wj = static_cast<T>(ck);
          Wk[j] = wj;
        }
        S[pk<T>(n, k, j)] = static_cast<T>(div_or_zero(static_cast<double>(wj), den));
      }
      if (j == k) {
        S[pk<T>(n, k, k)] = w;
        piv[k] = 1;
      }
      kk += 1;
    } else if (kind == -1) {
      // interchange k <-> r (:414-498), then the candidate row is row k
      const double den = safe_den(crr);
      if (j < k) {
        swap_elems(&S[pk<T>(n, j, k)], &S[pk<T>(n, j, r)]);
        if (j < kk) swap_elems(&W[j * n + k], &W[j * n + r]);
      } else if (j == k) {
        S[pk<T>(n, r, r)] = S[pk<T>(n, k, k)];
        S[pk<T>(n, k, k)] = static_cast<T>(crr);
        Wk[k] = static_cast<T>(crr);
        piv[k] = 1;
        const int t = perm[k];
        perm[k] = perm[r];
        perm[r] = t;
      } else if (j < n) {
        const T old = S[pk<T>(n, k, j)];
        if (j > r)
          S[pk<T>(n, r, j)] = old;
        else if (j < r)
          S[pk<T>(n, j, r)] = old;
        const T wj = j == r ? static_cast<T>(ck) : c;
        Wk[j] = wj;
        S[pk<T>(n, k, j)] = static_cast<T>(div_or_zero(static_cast<double>(wj), den));
      }
      kk += 1;
    } else {
      // 2x2 pivot at (k, k+1) after the interchange k+1 <-> r (:500-587),
      // ScaleTwoByTwoPivot (:658-704)
      const int k1 = k + 1;
      T* Wk1 = W + (kk + 1) * n;
      const bool sw = r != k1;
      // W(k, k+1) after the interchange is the candidate's value at k (the
      // candidate row overwrote W(k, r), which the interchange moves to
      // k+1: :184-247, :510-514), and the D block
      const double a = wkk;
      const double b = ck;
      const double cc = crr;
      double itl, ibr, ioff;
      inverse_2x2_nc(a, b, cc, &itl, &ibr, &ioff);
      if (j < k) {
        if (sw) {
          swap_elems(&S[pk<T>(n, j, k1)], &S[pk<T>(n, j, r)]);
          if (j < kk) swap_elems(&W[j * n + k1], &W[j * n + r]);
        }
      } else if (j == k) {
        S[pk<T>(n, k, k)] = static_cast<T>(a);
        S[pk<T>(n, k, k1)] = static_cast<T>(b);
        Wk[k1] = static_cast<T>(b);
        piv[k] = 2;
      } else if (j == k1) {
        if (sw) {
          S[pk<T>(n, r, r)] = S[pk<T>(n, k1, k1)];
          const int t = perm[k1];
          perm[k1] = perm[r];
          perm[r] = t;
        }
        S[pk<T>(n, k1, k1)] = static_cast<T>(cc);
        Wk1[k1] = static_cast<T>(cc);
        piv[k1] = -2;
      } else if (j < n) {
        if (sw) {
          const T old = S[pk<T>(n, k1, j)];
          if (j > r)
            S[pk<T>(n, r, j)] = old;
          else if (j < r)
            S[pk<T>(n, j, r)] = old;
        }
        const T w0 = (sw && j == r) ? wk_k1 : w;
        const T w1 = (sw && j == r) ? static_cast<T>(ck1) : c;
        Wk[j] = w0;
        Wk1[j] = w1;
        double u0, u1;
        apply_2x2_nc(itl, ibr, ioff, static_cast<double>(w0), static_cast<double>(w1), &u0, &u1);
        S[pk<T>(n, k, j)] = static_cast<T>(u0);
        S[pk<T>(n, k1, j)] = static_cast<T>(u1);
      }
      kk += 2;
    }
    __syncthreads();
  }
  return e;
}

// ---- resident mode (batches of matrices up to kBatchResidentN) -------------
// Every panel of one matrix, then its trailing update (the reference's
// blocked loop, src/block_ldlt_decomposition.cc:50-59), on a storage view:
// the fused on-chip panel for a matrix held whole in LDS, factor_panel for
// one that keeps its first rows in global memory.
template <typename T, typename Mat>
__device__ void factor_resident(const Mat& F, int n, int* perm, int* piv, int bw, int ts, T* W,
                                T* Lk, T* Lr, ScaleTab& st, double* rv, int* ri,
                                PivotSlots& ps) {
  constexpr bool kOnChip = std::is_same<Mat, LdsMat<T>>::value;
  int s = 0;
  int par = 0;
  for (int P = 0; s < n; ++P) {
    int e;
    if constexpr (kOnChip)
      e = factor_panel_lds<T, sizeof(T) == 8>(F.s, n, perm, piv, bw, P, s, W, ps, rv, ri, par);
    else
      e = factor_panel<T, 256, sizeof(T) == 8>(F, n, perm, piv, bw, P, s, W, Lk, Lr, st, rv, ri);
    if (ts == 32)
      mfma_update<T, 256, 32>(F, n, s, e, W);
    else
      mfma_update<T, 256, 16>(F, n, s, e, W);
    __syncthreads();
    s = e;
  }
}

// Packed index of element (r, c), r <= c, of an order-m upper triangle.
__device__ __forceinline__ int packed_at(int m, int r, int c) {
  return r * m - ((r * (r - 1)) >> 1) + (c - r);
}

// Copy the upper triangle of rows/columns [R0, n) between A (row-major,
// lda) and the packed LDS array S (m = n - R0): 16 elements per thread in
// flight, consecutive threads on consecutive columns of a row.
template <typename T, bool TO_LDS>
__device__ void packed_copy(T* __restrict__ A, int64_t lda, int R0, int m, T* S) {
  const int total = m * (m + 1) / 2;
  // thread-local walk over the packed index idx = tid + 256 u: row r holds
  // [start, start + m - r)
  int r = 0, start = 0;
  for (int base = 0; base < total; base += 256 * 16) {
    int rr[16], cc[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int idx = base + 256 * u + static_cast<int>(threadIdx.x);
      if (idx < total) {
        while (idx >= start + (m - r)) {
          start += m - r;
          ++r;
        }
        rr[u] = r;
        cc[u] = r + (idx - start);
      } else {
        rr[u] = -1;
        cc[u] = 0;
      }
    }
    if constexpr (TO_LDS) {
      T v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u)
        v[u] = rr[u] >= 0 ? A[(int64_t)(R0 + rr[u]) * lda + R0 + cc[u]] : T(0);
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (rr[u] >= 0) S[packed_at(m, rr[u], cc[u])] = v[u];
    } else {
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (rr[u] >= 0) A[(int64_t)(R0 + rr[u]) * lda + R0 + cc[u]] = S[packed_at(m, rr[u], cc[u])];
    }
  }
}

// Interchanges deferred by a HybridMat (dlist: ns pairs, in order) applied
// to rows [0, R0) of A, composed into one gather per row (compose_swaps;
// slot and content: LDS scratch of 2 ns ints each).  Eight rows per pass,
// every read of a pass before its writes.
template <typename T>
__device__ void apply_deferred(T* __restrict__ A, int64_t lda, int R0, const int* dlist, int ns,
                               int* slot, int* content) {
  if (ns == 0 || R0 == 0) return;
  compose_swaps(dlist, ns, slot, content);
  const int ne = 2 * ns;
  const int t0 = threadIdx.x, t1 = threadIdx.x + 256;
  const int d0 = t0 < ne ? dlist[t0] : 0, s0 = t0 < ne ? content[slot[t0]] : 0;
  const int d1 = t1 < ne ? dlist[t1] : 0, s1 = t1 < ne ? content[slot[t1]] : 0;
  for (int p0 = 0; p0 < R0; p0 += 8) {
    T v0[8], v1[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int row = p0 + u;
      const T* R = A + (int64_t)row * lda;
      v0[u] = (row < R0 && t0 < ne) ? R[s0] : T(0);
      v1[u] = (row < R0 && t1 < ne) ? R[s1] : T(0);
    }
    __syncthreads();  // every read of these rows precedes every write
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int row = p0 + u;
      T* R = A + (int64_t)row * lda;
      if (row < R0 && t0 < ne) R[d0] = v0[u];
      if (row < R0 && t1 < ne) R[d1] = v1[u];
    }
    __syncthreads();
  }
}

// Resident persistent kernel: workgroups take matrices from the work queue
// (largest first); per matrix the dynamic LDS holds, in order, the panel
// rows W ((bw + 2) x n), perm and piv (n ints each), the deferred
// interchange list (2 n ints) and the packed upper triangle of the trailing
// square [R0, n)^2, R0 the smallest start whose triangle fits what is left
// (R0 = 0 for every matrix that fits whole: LdsMat, no global access until
// the results are written back).
template <typename T>
__global__ __launch_bounds__(256) void k_batched_lds(BatchArgs<T> a) {
  __shared__ T Lk[kBatchMaxWidth + 2];
  __shared__ T Lr[kBatchMaxWidth + 2];
  __shared__ ScaleTab st;
  __shared__ double rv[8];  // two halves: the one-barrier reductions alternate
  __shared__ int ri[8];
  __shared__ PivotSlots ps;
  __shared__ int sq;
  __shared__ int dcount;
  unsigned char* arena = bbk_batched_lds;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) sq = atomicAdd(a.queue, 1);
    __syncthreads();
    const int q = sq;
    __syncthreads();
    if (q >= a.count) break;
    const int mi = a.order[q];
    const int n = a.ns[mi];
    if (n <= 0) continue;
    const int c = batch_class(n);
    const int bw = a.widths.w[c];
    const int ts = a.widths.ts[c] == 32 ? 32 : 16;
    T* A = a.A[mi];
    const int64_t lda = a.lda[mi];
    // LDS layout (16-byte aligned regions)
    // (the W region also holds the 4 n ints of apply_deferred's scratch)
    T* W = reinterpret_cast<T*>(arena);
    size_t off = (std::max((size_t)(bw + 2) * n * sizeof(T), (size_t)16 * n) + 15) / 16 * 16;
    int* lperm = reinterpret_cast<int*>(arena + off);
    int* lpiv = lperm + n;
    int* dlist = lpiv + n;
    off += ((size_t)4 * n * sizeof(int) + 15) / 16 * 16;
    T* S = reinterpret_cast<T*>(arena + off);
    const int avail = static_cast<int>((a.lds_bytes - (int)off) / (int)sizeof(T));
    int m = n;
    while (m > 0 && m * (m + 1) / 2 > avail) --m;
    const int R0 = n - m;
    for (int i = tid; i < n; i += 256) {
      lperm[i] = i;
      lpiv[i] = 1;
    }
    if (tid == 0) dcount = 0;
    packed_copy<T, true>(A, lda, R0, m, S);
    __syncthreads();
    if (R0 == 0)
      factor_resident<T>(LdsMat<T>{S, n}, n, lperm, lpiv, bw, ts, W, Lk, Lr, st, rv, ri, ps);
    else
      factor_resident<T>(HybridMat<T>{A, lda, S, n, R0, dlist, &dcount}, n, lperm, lpiv, bw, ts,
                         W, Lk, Lr, st, rv, ri, ps);
    // results: the trailing square, perm and piv back to global memory;
    // the deferred interchanges on the rows kept there
    packed_copy<T, false>(A, lda, R0, m, S);
    int* perm = a.perm[mi];
    int* piv = a.piv[mi];
    for (int i = tid; i < n; i += 256) {
      perm[i] = lperm[i];
      piv[i] = lpiv[i];
    }
    // W is free now: slot and content of the composed interchanges
    int* slot = reinterpret_cast<int*>(W);
    apply_deferred<T>(A, lda, R0, dlist, R0 > 0 ? dcount : 0, slot, slot + 2 * n);
    __syncthreads();
  }
}

// ---- symmetric-workspace mode (one wavefront per matrix, n <= 256) ---------
// Each workgroup is one wavefront and owns a slice of the batch workspace
// holding a full symmetric copy S of the matrix it factors (both triangles,
// row stride n).  An interchange moves no data: the logical -> physical map
// pi is updated instead, and logical element (i, j) is S(pi[i], pi[j]).
//  * Every row the panel reads -- the pivot row and the candidate row, whose
//    column part (A(i, r), i < r) is a strided walk in upper storage -- is one
//    physical row of S, read with a gather that is contiguous where pi is.
//  * The reference's column swaps of the finished rows
//    (src/block_ldlt_decomposition.cc:414-587) cost nothing: a finished
//    row's values stay at the physical columns they were written to, and
//    its final column j is S(pi[k], pi[j]) with the final map.
//  * The pivot loop stores nothing to global memory.  Rows k + 1 and k + 2
//    are loaded one pivot ahead and, when the interchange touches them,
//    patched in registers (S is symmetric, so the new entry of row x at
//    column r is row k's raw value at column x); the candidate row is the
//    one load a pivot waits for.  The scaled rows U and the D blocks are
//    written once per panel, rebuilt from W and the per-row scaling record
//    (bit for bit the values the scaling forms), and the trailing update
//    (MFMA tiles over the logical upper triangle, each result stored at its
//    place and its mirror) takes its U operands from the same rebuild.
// Positions: lane l holds columns l + 64 u (u < kSymNP) of every row; pi of
// those columns sits in registers (pj), and in LDS for the update.  The
// arithmetic is factor_panel's (same row products, decisions, W updates and
// scaling, src/block_ldlt_decomposition.cc:80-120,122-704) and
// mfma_update's, element for element, so the factors equal the
// global-memory team's with MFMA update tiles bit for bit.  Loads are issued
// from clamped (always valid) addresses and their values masked after, so
// no load sits in a branch of its own.
constexpr int kSymNP = 4;         // positions per lane: orders up to 256
constexpr int kSymMaxWidth = 32;  // widest panel of this mode (register budget)

// The matrix pointers come from a device array, so the compiler cannot
// tell they are global; this view makes their accesses global loads and
// stores (flat ones would also wait on the LDS counter).
template <typename T>
using gmem_t = __attribute__((address_space(1))) T;
template <typename T>
__device__ __forceinline__ gmem_t<T>* to_global(T* p) {
  return (gmem_t<T>*)p;
}

template <typename T>
__device__ __forceinline__ T readlane_t(T v, int lane) {
  if constexpr (sizeof(T) == 8) {
    return readlane_f64(v, lane);
  } else {
    static_assert(sizeof(T) == 4, "readlane of a 4-byte value");
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
  }
}

// The value at position j (0 <= j < 64 kSymNP, wave-uniform) of a per-lane
// position array whose lane l, slot u holds position l + 64 u.  pos_get
// reads another lane's register, so it is called where every lane is
// active.  (Constant subscripts only, so the arrays stay in registers.)
static_assert(kSymNP == 4, "pos_get / pos_set select among four slots");
template <typename T>
__device__ __forceinline__ T pos_get(const T (&v)[kSymNP], int j) {
  const int u = j >> 6;
  const T x = u == 0 ? v[0] : u == 1 ? v[1] : u == 2 ? v[2] : v[3];
  return readlane_t(x, j & 63);
}
template <typename T>
__device__ __forceinline__ void pos_set(T (&v)[kSymNP], int j, T val) {
  const bool mine = (threadIdx.x & 63) == (j & 63);
  const int u = j >> 6;
  v[0] = (mine && u == 0) ? val : v[0];
  v[1] = (mine && u == 1) ? val : v[1];
  v[2] = (mine && u == 2) ? val : v[2];
  v[3] = (mine && u == 3) ? val : v[3];
}

// Logical row x at every position (values at positions past n unused).
template <typename T>
__device__ __forceinline__ void sym_row(const gmem_t<T>* S, int n, const int (&pj)[kSymNP], int x,
                                        T (&v)[kSymNP]) {
  const gmem_t<T>* row = S + static_cast<unsigned>(pos_get(pj, x) * n);
#pragma unroll
  for (int u = 0; u < kSymNP; ++u) v[u] = row[static_cast<unsigned>(pj[u])];
}

// Row x, loaded before this pivot's interchange of ps and r, made current:
// x == r: the row now lives where row ps lived, so it is row ps (src) with
// its column r taken from src's diagonal; otherwise only column r changed,
// to S(pi_old[ps], pi[x]), which by symmetry is src at column x.
template <typename T>
__device__ __forceinline__ void sym_patch(T (&rx)[kSymNP], int x, const T (&src)[kSymNP], int ps,
                                          int r) {
  if (x == r) {
    const T d = pos_get(src, ps);
#pragma unroll
    for (int u = 0; u < kSymNP; ++u) rx[u] = src[u];
    pos_set(rx, r, d);
  } else if (r > x) {
    pos_set(rx, r, pos_get(src, x));
  }
}

// Panel P of one matrix held in S (see above): the pivot loop; returns the
// panel's end e.  pj: pi of the lane's positions (updated by the
// interchanges); spiv: the pivot types (LDS).
template <typename T, bool ACC4>
__device__ int sym_panel(const gmem_t<T>* S, int n, int (&pj)[kSymNP], int* spiv, int bw, int P,
                         int s, T* W, T* Lk, T* Lr, ScaleTab& st) {
  const int lane = threadIdx.x & 63;
  const int j0 = min(lane, n - 1);
  int e = min((P + 1) * bw, n);
  T rk0[kSymNP], rk1[kSymNP], rk2[kSymNP];
  sym_row<T>(S, n, pj, s, rk0);
#pragma unroll
  for (int u = 0; u < kSymNP; ++u) rk1[u] = rk2[u] = T(0);
  // nothing in flight when the pivot loop starts, so the compiler's waits
  // inside it count only the loop's own loads (vmcnt(0); gfx9 encoding)
  __builtin_amdgcn_s_waitcnt(0x0F70);
  bool have1 = false;
  int kk = 0;
  while (s + kk < e) {
    const int k = s + kk;
    if (!have1 && k + 1 < n) sym_row<T>(S, n, pj, k + 1, rk1);
    if (k + 2 < n) sym_row<T>(S, n, pj, k + 2, rk2);
    if (lane < kk) Lk[lane] = ucol<T>(st, W, n, lane, k);  // L(k, s + lane)
    __syncthreads();
    // W(k, j) = A(k, j) - sum_p L(k, p) W(p, j), j >= k  (:122-162)
    T w[kSymNP], dot[kSymNP];
    panel_rows_dot_multi<T, ACC4, kSymNP, 64>(Lk, W, n, kk, j0, dot);
    T* Wk = W + (int64_t)kk * n;
    double best = 0.0;
    int arg = n;
#pragma unroll
    for (int u = 0; u < kSymNP; ++u) {
      const int j = lane + 64 * u;
      w[u] = rk0[u] - dot[u];
      if (j >= k && j < n) {
        Wk[j] = w[u];
        const double a = fabs(static_cast<double>(w[u]));
        if (j > k && a > best) {
          best = a;
          arg = j;
        }
      }
    }
    wave_argmax(best, arg);  // (:164-182) first maximum
    const double omega = best;
    const int r = arg;
    const double wkk = static_cast<double>(pos_get(w, k));
    const double akk = fabs(wkk);
    int kind = 1;
    T c[kSymNP];
    double ck = 0.0, ck1 = 0.0, crr = 0.0;
    if (!(akk >= kAlpha * omega || omega == 0.0)) {
      // candidate row r: C(i) = A(r, i) - sum_p L(r, p) W(p, i), i >= k
      // (:184-352), sigma over i != r (:354-367); A(r, i) = S(pi[r], pi[i])
      T cr[kSymNP];
      sym_row<T>(S, n, pj, r, cr);
      if (lane < kk) Lr[lane] = ucol<T>(st, W, n, lane, r);  // L(r, s + lane)
      __syncthreads();
      panel_rows_dot_multi<T, ACC4, kSymNP, 64>(Lr, W, n, kk, j0, dot);
      double sig = 0.0;
#pragma unroll
      for (int u = 0; u < kSymNP; ++u) {
        const int i = lane + 64 * u;
        c[u] = cr[u] - dot[u];
        const double a = fabs(static_cast<double>(c[u]));
        if (i >= k && i < n && i != r) sig = (sig < a) ? a : sig;
      }
      const double sigma = wave_max_nan_ignoring(sig);
      crr = static_cast<double>(pos_get(c, r));
      ck = static_cast<double>(pos_get(c, k));
      ck1 = k + 1 < n ? static_cast<double>(pos_get(c, k + 1)) : 0.0;
      kind = select_pivot_nc(akk, omega, sigma, fabs(crr));
      if (kind == 2 && kk == e - s - 1 && e < n) {
        // 2x2 across the panel edge: this panel ends before k (:589-631)
        e = k;
        __syncthreads();
        break;
      }
      if (kind == -1) {
        // interchange k <-> r (:414-498): the finished rows' W columns, pi;
        // row k of W becomes the candidate row with W(k, r) = C(k)
        // (the candidate's value at k, :184-247) and W(k, k) = C(r)
        if (lane < kk) swap_elems(&W[(int64_t)lane * n + k], &W[(int64_t)lane * n + r]);
#pragma unroll
        for (int u = 0; u < kSymNP; ++u) {
          const int j = lane + 64 * u;
          if (j > k && j < n && j != r) Wk[j] = c[u];
        }
        if (lane == 0) {
          Wk[k] = static_cast<T>(crr);
          Wk[r] = static_cast<T>(ck);
        }
        const int pk = pos_get(pj, k), pr = pos_get(pj, r);
        pos_set(pj, k, pr);
        pos_set(pj, r, pk);
        sym_patch<T>(rk1, k + 1, rk0, k, r);
        sym_patch<T>(rk2, k + 2, rk0, k, r);
      } else if (kind == 2) {
        // interchange k+1 <-> r (:500-587); row k+1 of W becomes the
        // candidate row; W(k, k+1) = C(k), W(k, r) = the old W(k, k+1)
        const int k1 = k + 1;
        T* Wk1 = W + (int64_t)(kk + 1) * n;
        const T wk_k1 = pos_get(w, k1);
        if (r != k1 && lane < kk)
          swap_elems(&W[(int64_t)lane * n + k1], &W[(int64_t)lane * n + r]);
#pragma unroll
        for (int u = 0; u < kSymNP; ++u) {
          const int j = lane + 64 * u;
          if (j >= k + 2 && j < n && j != r) Wk1[j] = c[u];
        }
        if (lane == 0) {
          Wk[k1] = static_cast<T>(ck);
          if (r != k1) {
            Wk[r] = wk_k1;
            Wk1[r] = static_cast<T>(ck1);
          }
          Wk1[k1] = static_cast<T>(crr);
        }
        if (r != k1) {
          const int pa = pos_get(pj, k1), pb = pos_get(pj, r);
          pos_set(pj, k1, pb);
          pos_set(pj, r, pa);
          sym_patch<T>(rk2, k + 2, rk1, k1, r);
        }
      } else if (lane == 0) {
        Wk[r] = static_cast<T>(ck);  // (:184-247)
      }
    }
    if (kind != 2) {
      // ScaleOneByOnePivot (:633-656): the record; U and D are written at
      // the panel's end
      if (lane == 0) {
        spiv[k] = 1;
        st.ty[kk] = 1;
        st.c0[kk] = safe_den(kind == -1 ? crr : wkk);
      }
      kk += 1;
#pragma unroll
      for (int u = 0; u < kSymNP; ++u) {
        rk0[u] = rk1[u];
        rk1[u] = rk2[u];
      }
      have1 = true;
    } else {
      // ScaleTwoByTwoPivot (:658-704), binary64 without contraction:
      // D = [W(k,k) C(k); C(k) C(r)]
      if (lane == 0) {
        double itl, ibr, ioff;
        inverse_2x2_nc(wkk, ck, crr, &itl, &ibr, &ioff);
        spiv[k] = 2;
        spiv[k + 1] = -2;
        st.ty[kk] = 2;
        st.c0[kk] = itl;
        st.c1[kk] = ioff;
        st.ty[kk + 1] = -2;
        st.c0[kk + 1] = ioff;
        st.c1[kk + 1] = ibr;
      }
      kk += 2;
#pragma unroll
      for (int u = 0; u < kSymNP; ++u) rk0[u] = rk2[u];
      have1 = false;
    }
    __syncthreads();
  }
  return e;
}

// The panel's finished rows s..e-1 into S: row s+p at physical row
// pi[s+p]: D(p, p) = W(p, s+p), D(p, p+1) = W(p, s+p+1) for the first row
// of a 2x2, and U(p, j) = ucol(p, j) at every later column (the values
// ScaleOneByOnePivot / ScaleTwoByTwoPivot form, :633-704).
template <typename T>
__device__ void sym_write_rows(gmem_t<T>* S, int n, const int (&pj)[kSymNP], int s, int e,
                               const T* W, const ScaleTab& st) {
  const int lane = threadIdx.x & 63;
  for (int p = 0; p < e - s; ++p) {
    const int k = s + p;
    gmem_t<T>* row = S + static_cast<unsigned>(pos_get(pj, k) * n);
    const int dlast = st.ty[p] == 2 ? k + 1 : k;  // last column holding D
#pragma unroll
    for (int u = 0; u < kSymNP; ++u) {
      const int j = lane + 64 * u;
      if (j >= k && j < n) {
        const int jc = j;
        row[static_cast<unsigned>(pj[u])] = j <= dlast ? W[p * n + jc] : ucol<T>(st, W, n, p, jc);
      }
    }
  }
}

// Trailing update of the logical upper triangle of [e, n)^2,
// C(i, j) -= sum_{p < kp} U(p, i) W(p, j), as mfma_update does it (same K
// steps of 4, U negated, accumulators started from C), each result stored
// at (i, j) and at its mirror.  U comes from ucol (LDS), C from S; a
// wavefront walks the tile rows, loading the next tile's C values (two
// register sets, used in turn) while the current tile's MFMAs run.
// Operands of rows i >= n or columns j >= n only reach outputs that are
// not stored, so only the K rows past kp (in the last, partial K step) are
// masked.  pi: LDS copy of the map.
template <typename T, int TS>
__device__ void sym_update(gmem_t<T>* S, int n, const int* pi, int s, int e, const T* W,
                           const ScaleTab& st) {
  constexpr int TB = TS / 16;
  constexpr int KS = (kSymMaxWidth + 1 + 3) / 4;
  typedef typename Mfma<T>::acc_t acc_t;
  const int lane = threadIdx.x & 63;
  const int m = n - e, kp = e - s;
  if (m <= 0 || kp <= 0) return;
  const int q = lane >> 4, cl = lane & 15;
  const int nkf = kp >> 2;            // K steps whose four rows are all < kp
  const int nk = (kp + 3) >> 2;       // with the partial one
  const int nt = (m + TS - 1) / TS;
  for (int I = 0; I < nt; ++I) {
    const int i0 = e + I * TS;
    // physical rows of the lane's outputs (and those times n): every S
    // offset fits 32 bits (n <= 256)
    unsigned prow[TB][4], prn[TB][4];
#pragma unroll
    for (int ti = 0; ti < TB; ++ti)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        prow[ti][v] = pi[min(i0 + 16 * ti + Mfma<T>::row(lane, v), n - 1)];
        prn[ti][v] = prow[ti][v] * n;
      }
    // -U(p, i) for the tile row's K steps
    T a[TB][KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int p = 4 * ks + q;
#pragma unroll
      for (int tb = 0; tb < TB; ++tb) {
        const int ia = min(i0 + 16 * tb + cl, n - 1);
        if (ks < nkf) {
          a[tb][ks] = -ucol<T>(st, W, n, p, ia);
        } else if (ks < nk) {
          const T u = ucol<T>(st, W, n, min(p, kp - 1), ia);
          a[tb][ks] = p < kp ? -u : T(0);
        } else {
          a[tb][ks] = T(0);
        }
      }
    }
    auto load_c = [&](int J, acc_t (&acc)[TB][TB], unsigned (&pc)[TB]) {
      const int jj = e + J * TS;
#pragma unroll
      for (int tj = 0; tj < TB; ++tj) pc[tj] = pi[min(jj + 16 * tj + cl, n - 1)];
#pragma unroll
      for (int ti = 0; ti < TB; ++ti)
#pragma unroll
        for (int tj = 0; tj < TB; ++tj)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[ti][tj][v] = S[prn[ti][v] + pc[tj]];
    };
    auto tile = [&](int J, acc_t (&acc)[TB][TB], const unsigned (&pc)[TB]) {
      const int jj = e + J * TS;
      // W(p, j) for every K step of the tile, read together
      T b[KS][TB];
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int p = 4 * ks + q;
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) {
          const int jb = min(jj + 16 * tb + cl, n - 1);
          if (ks < nkf) {
            b[ks][tb] = W[p * n + jb];
          } else if (ks < nk) {
            const T wv = W[min(p, kp - 1) * n + jb];
            b[ks][tb] = p < kp ? wv : T(0);
          } else {
            b[ks][tb] = T(0);
          }
        }
      }
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        if (ks < nk) {
#pragma unroll
          for (int ti = 0; ti < TB; ++ti)
#pragma unroll
            for (int tj = 0; tj < TB; ++tj)
              acc[ti][tj] = Mfma<T>::mma(a[ti][ks], b[ks][tj], acc[ti][tj]);
        }
      }
#pragma unroll
      for (int tj = 0; tj < TB; ++tj) {
        const unsigned pcn = pc[tj] * n;
#pragma unroll
        for (int ti = 0; ti < TB; ++ti)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = i0 + 16 * ti + Mfma<T>::row(lane, v), j = jj + 16 * tj + cl;
            if (i <= j && j < n) {
              S[prn[ti][v] + pc[tj]] = acc[ti][tj][v];
              if (i < j) S[pcn + prow[ti][v]] = acc[ti][tj][v];
            }
          }
      }
    };
    acc_t c0[TB][TB], c1[TB][TB];
    unsigned pc0[TB], pc1[TB];
    load_c(I, c0, pc0);
    for (int J = I; J < nt; J += 2) {
      const bool two = J + 1 < nt;
      if (two) load_c(J + 1, c1, pc1);
      tile(J, c0, pc0);
      if (two) {
        if (J + 2 < nt) load_c(J + 2, c0, pc0);
        tile(J + 1, c1, pc1);
      }
    }
  }
}

// S := the symmetric matrix whose upper triangle is A's, R rows at a time
// (R in {4, 8, 16}, four rows in flight per pass).  Each row goes to S as
// it is and to a copy in LDS (stage, R x (n + 1) elements, a stride that
// keeps the transposed reads off one bank); the strictly lower half is
// then written from the copy, R consecutive elements of an S row per
// lane group, instead of one element per row.  R = 0: no LDS, each
// element's mirror stored on its own.
template <typename T>
__device__ void sym_load(const gmem_t<T>* A, int64_t lda, int n, gmem_t<T>* S, T* stage,
                         int R) {
  const int lane = threadIdx.x & 63;
  const int ld = n + 1;
  const int rows = R > 0 ? R : 4;
  for (int i0 = 0; i0 < n; i0 += rows) {
    for (int d0 = 0; d0 < rows; d0 += 4) {
      T v[4][kSymNP];
#pragma unroll
      for (int d = 0; d < 4; ++d)
#pragma unroll
        for (int u = 0; u < kSymNP; ++u) {
          const int i = min(i0 + d0 + d, n - 1), j = min(i + lane + 64 * u, n - 1);
          v[d][u] = A[i * lda + static_cast<unsigned>(j)];
        }
#pragma unroll
      for (int d = 0; d < 4; ++d)
#pragma unroll
        for (int u = 0; u < kSymNP; ++u) {
          const int i = i0 + d0 + d, j = i + lane + 64 * u;
          if (i < n && j < n) {
            S[static_cast<unsigned>(i * n + j)] = v[d][u];
            if (R > 0)
              stage[(d0 + d) * ld + j] = v[d][u];
            else
              S[static_cast<unsigned>(j * n + i)] = v[d][u];
          }
        }
    }
    if (R > 0) {
      __syncthreads();
      // S(j, i) for i in [i0, i0 + R), i < j < n: lane group jr takes row
      // j, lane ic column i0 + ic
      const int per = 64 / R, ic = lane & (R - 1), jr = lane / R;
      for (int j0 = i0 + 1; j0 < n; j0 += 4 * per) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int j = j0 + t * per + jr, i = i0 + ic;
          if (j < n && i < j) S[static_cast<unsigned>(j * n + i)] = stage[ic * ld + j];
        }
      }
      __syncthreads();
    }
  }
}

// Rows per pass of sym_load for order n when `elems` elements of LDS are
// free: the largest of 16, 8, 4 whose R x (n + 1) copy fits, else 0.
__device__ __forceinline__ int sym_load_rows(int n, int elems) {
  for (int R = 16; R >= 4; R >>= 1)
    if (R * (n + 1) <= elems) return R;
  return 0;
}

// A(k, j) := S(pi[k], pi[j]) for k <= j (the factors in the final logical
// order), perm := pi, piv from LDS; four rows in flight per pass.
template <typename T>
__device__ void sym_store(const gmem_t<T>* S, int n, const int* pi, const int* spiv,
                          gmem_t<T>* A, int64_t lda, gmem_t<int>* perm, gmem_t<int>* piv) {
  const int lane = threadIdx.x & 63;
  int pj[kSymNP];
#pragma unroll
  for (int u = 0; u < kSymNP; ++u) pj[u] = pi[min(lane + 64 * u, n - 1)];
  for (int k0 = 0; k0 < n; k0 += 4) {
    T v[4][kSymNP];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const gmem_t<T>* row = S + static_cast<unsigned>(pi[min(k0 + d, n - 1)] * n);
#pragma unroll
      for (int u = 0; u < kSymNP; ++u) v[d][u] = row[static_cast<unsigned>(pj[u])];
    }
#pragma unroll
    for (int d = 0; d < 4; ++d)
#pragma unroll
      for (int u = 0; u < kSymNP; ++u) {
        const int k = k0 + d, j = lane + 64 * u;
        if (k < n && j >= k && j < n) A[k * lda + static_cast<unsigned>(j)] = v[d][u];
      }
  }
  for (int i = lane; i < n; i += 64) {
    perm[i] = pi[i];
    piv[i] = spiv[i];
  }
}

// Persistent kernel of the symmetric-workspace mode: workgroups (one
// wavefront each) take matrices from the work queue (largest first).
// Dynamic LDS per matrix of order n and panel width bw: W ((bw + 1) x n),
// pi and the pivot types (n ints each).  Workspace slice of the workgroup:
// n x n elements of S.
// (binary64: LDS admits one wavefront per SIMD, so it may take all 512
// registers; binary32 keeps two per SIMD)
template <typename T>
__global__ __launch_bounds__(64, sizeof(T) == 8 ? 1 : 2) void k_batched_sym(BatchArgs<T> a) {
  __shared__ T Lk[kSymMaxWidth + 2];
  __shared__ T Lr[kSymMaxWidth + 2];
  __shared__ ScaleTab st;
  __shared__ int sq;
  T* lds = reinterpret_cast<T*>(bbk_batched_lds);
  gmem_t<T>* S = to_global(a.ws) + (int64_t)blockIdx.x * a.ws_stride;
  const int lane = threadIdx.x & 63;
  for (;;) {
    if (lane == 0) sq = atomicAdd(a.queue, 1);
    __syncthreads();
    const int q = sq;
    __syncthreads();
    if (q >= a.count) break;
    const int mi = a.order[q];
    const int n = a.ns[mi];
    if (n <= 0) continue;
    const int c = batch_class(n);
    const int bw = a.widths.w[c];
    T* W = lds;
    int* pi = reinterpret_cast<int*>(W + (int64_t)(bw + 1) * n);
    int* spiv = pi + n;
    gmem_t<T>* A = to_global(a.A[mi]);
    const int64_t lda = a.lda[mi];
    int pj[kSymNP];
#pragma unroll
    for (int u = 0; u < kSymNP; ++u) pj[u] = min(lane + 64 * u, n - 1);
    // the W rows' LDS ((bw + 1) x n elements) is free until the first panel
    sym_load<T>(A, lda, n, S, W, sym_load_rows(n, (bw + 1) * n));
    const int ts = a.widths.ts[c] == 32 ? 32 : 16;
    int s = 0;
    for (int P = 0; s < n; ++P) {
      const int e = sym_panel<T, sizeof(T) == 8>(S, n, pj, spiv, bw, P, s, W, Lk, Lr, st);
#pragma unroll
      for (int u = 0; u < kSymNP; ++u)
        if (lane + 64 * u < n) pi[lane + 64 * u] = pj[u];
      __syncthreads();
      sym_write_rows<T>(S, n, pj, s, e, W, st);
      if (ts == 32)
        sym_update<T, 32>(S, n, pi, s, e, W, st);
      else
        sym_update<T, 16>(S, n, pi, s, e, W, st);
      __syncthreads();
      s = e;
    }
    sym_store<T>(S, n, pi, spiv, A, lda, to_global(a.perm[mi]), to_global(a.piv[mi]));
    __syncthreads();
  }
}

// Dynamic LDS bytes of k_batched_sym for orders up to nmax, widths up to wmax.
__host__ __device__ constexpr int sym_lds_bytes(int nmax, int wmax, int elem_bytes) {
  return ((wmax + 1) * nmax * elem_bytes + 2 * nmax * 4 + 15) / 16 * 16;
}

// ---- lockstep mode (batches of large matrices) -----------------------------
// Per-matrix state (4 ints): next panel start, this panel's start, this
// panel's rows (0: nothing to update), unused.
template <typename T>
struct StepArgs {
  int count;
  int step;
  const int* ns;
  T* const* A;
  const int64_t* lda;
  int* const* perm;
  int* const* piv;
  T* ws;                 // W slices, (width + 2) x n elements per matrix
  const int64_t* wsoff;  // matrix -> first element of its slice
  int* st;
  const int* prefix;     // this step's first tile per matrix (count + 1 entries)
  BatchWidths widths;
  int lds_dyn;           // dynamic LDS bytes of k_step_panel<T, true>
  const int* order;      // panel workgroup -> matrix (largest first)
};

// Rows [0, R) of A: row[dst[t]] = row[src[t]] for t < ne (one composed
// interchange gather, ne <= 192).  Each wavefront takes whole rows, lane l
// moving entries l, l + 64 and l + 128, four rows per pass; every load of
// a pass completes (wait_vm) before its stores, so a row's reads precede
// its writes; different rows are independent.
template <typename T>
__device__ void gather_rows(T* __restrict__ A, int64_t lda, int R, const int* dst,
                            const int* src, int ne) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int d[3], sc[3];
  bool h[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int t = lane + 64 * e;
    h[e] = t < ne;
    d[e] = h[e] ? dst[t] : 0;
    sc[e] = h[e] ? src[t] : 0;
  }
  for (int i0 = w; i0 < R; i0 += 4 * nw) {
    T v[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int row = i0 + u * nw;
      const T* Rr = A + (int64_t)(row < R ? row : 0) * lda;
#pragma unroll
      for (int e = 0; e < 3; ++e) v[u][e] = (row < R && h[e]) ? Rr[sc[e]] : T(0);
    }
    wait_vm();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int row = i0 + u * nw;
      if (row < R) {
        T* Rr = A + (int64_t)row * lda;
#pragma unroll
        for (int e = 0; e < 3; ++e)
          if (h[e]) Rr[d[e]] = v[u][e];
      }
    }
  }
}

// A panel's deferred interchanges (ds.list, ds.count pairs in order)
// applied to the finished rows [0, R) above it: composed into one gather
// (compose_swaps), then gather_rows.  Called by every thread of the team.
template <typename T, int NT>
__device__ void apply_panel_deferred(T* A, int64_t lda, int R, DeferScratch& ds) {
  static_assert(2 * DeferScratch::kMax <= 192, "gather_rows moves at most 192 entries");
  const int nd = ds.count;
  if (nd == 0 || R == 0) return;
  compose_swaps(ds.list, nd, ds.slot, ds.content);
  const int ne = 2 * nd;
  for (int t = threadIdx.x; t < ne; t += NT) ds.src[t] = ds.content[ds.slot[t]];
  __syncthreads();
  gather_rows<T>(A, lda, R, ds.list, ds.src, ne);
}

// Dynamic LDS limit of a lockstep panel's pivot and candidate rows (2 n
// elements); larger orders keep both rows in the global W slice.
constexpr size_t kStepRowLdsBytes = 60 * 1024;

// Positions per thread whose row products a lockstep panel forms together
// (W in a global slice: more loads in flight per thread).  Measured on the
// 1024-matrix batch (n in [256, 2048]) against one position at a time:
// binary32 113.7 -> 106.7 ms with 8 (2: 108.4, 4: 143.5), binary64
// 149.2 -> 144.0 ms with 4 (8: 210.2).
template <typename T>
constexpr int kStepPanelBatch = sizeof(T) == 4 ? 8 : 4;

// Interchanges of the rows above a lockstep panel deferred to one gather
// per row after the panel (binary32: 100.3 -> 88.7 ms on that batch;
// binary64, 16 elements per cache line instead of 32: 137.5 -> 139.4 ms,
// so it keeps the immediate column swaps).  (The same deferral in the
// persistent small-matrix kernel measured slower: 2.63 -> 3.28 ms.)
template <typename T>
constexpr bool kStepDefer = sizeof(T) == 4;

// One lockstep panel of matrix m: binary32 with deferred interchanges of
// the rows above the panel (applied before returning), binary64 with
// immediate ones.
template <typename T, bool LROW>
__device__ int step_panel(T* A, int64_t lda, int n, int* perm, int* piv, int bw, int P, int s,
                          T* W, T* Lk, T* Lr, ScaleTab& st, double* rv, int* ri, T* rowk,
                          T* rowc, T* Wl, int RL) {
  if constexpr (kStepDefer<T>) {
    __shared__ DeferScratch ds;
    if (threadIdx.x == 0) ds.count = 0;
    __syncthreads();
    const int e = factor_panel<T, 256, true, kStepPanelBatch<T>, LROW>(
        DeferGlobalMat<T>{A, lda, &ds}, n, perm, piv, bw, P, s, W, Lk, Lr, st, rv, ri, rowk,
        rowc, Wl, RL);
    __syncthreads();
    apply_panel_deferred<T, 256>(A, lda, s, ds);
    return e;
  } else {
    return factor_panel<T, 256, true, kStepPanelBatch<T>, LROW>(
        GlobalMat<T>{A, lda}, n, perm, piv, bw, P, s, W, Lk, Lr, st, rv, ri, rowk, rowc, Wl,
        RL);
  }
}

template <typename T, bool LROW>
__global__ __launch_bounds__(256) void k_step_panel(StepArgs<T> a) {
  __shared__ T Lk[kBatchMaxWidth + 2];
  __shared__ T Lr[kBatchMaxWidth + 2];
  __shared__ ScaleTab stab;
  __shared__ double rv[4];
  __shared__ int ri[4];
  const int m = a.order[blockIdx.x];
  const int n = a.ns[m];
  int* S = a.st + 4 * m;
  if (n <= 0) {
    if (threadIdx.x == 0) S[2] = 0;
    return;
  }
  const int bw = a.widths.w[batch_class(n)];
  const int P = a.step;
  if (P >= (n + bw - 1) / bw) {
    if (threadIdx.x == 0) S[2] = 0;
    return;
  }
  int* perm = a.perm[m];
  int* piv = a.piv[m];
  if (P == 0) {
    for (int i = threadIdx.x; i < n; i += 256) {
      perm[i] = i;
      piv[i] = 1;
    }
    __syncthreads();
  }
  const int s = P == 0 ? 0 : S[0];
  T* W = a.ws + a.wsoff[m];
  // LROW: the pivot row and the candidate row in dynamic LDS (n each),
  // then W's first RL rows (a multiple of 4, at most bw) in what is left
  T* rowk = reinterpret_cast<T*>(bbk_batched_lds);
  T* rowc = rowk + n;
  T* Wl = rowc + n;
  int RL = 0;
  if (LROW) {
    const int spare = a.lds_dyn / static_cast<int>(sizeof(T)) - 2 * n;
    RL = spare > 0 ? min(bw, spare / n) & ~3 : 0;
  }
  const int e = step_panel<T, LROW>(a.A[m], a.lda[m], n, perm, piv, bw, P, s, W, Lk, Lr, stab,
                                    rv, ri, rowk, rowc, Wl, RL);
  if (threadIdx.x == 0) {
    S[0] = e;
    S[1] = s;
    S[2] = e - s;
  }
}

template <typename T>
__global__ __launch_bounds__(256, 4) void k_step_update(StepArgs<T> a) {
  constexpr int TM = 64, KC = 8;
  __shared__ T Us[2][KC * (TM + 4)];
  __shared__ T Ws[2][KC * (TM + 4)];
  // workgroups are dealt to the 8 XCDs in turn (blockIdx mod 8); each XCD
  // takes one contiguous range of tiles instead, so the tiles of a matrix
  // share its U and W rows in one L2 (a bijection onto [0, gridDim.x))
  const int G = gridDim.x, q = G >> 3, r = G & 7;
  const int xcd = blockIdx.x & 7;
  const int b = xcd * q + min(xcd, r) + (blockIdx.x >> 3);
  // matrix m with prefix[m] <= b < prefix[m + 1]
  int lo = 0, hi = a.count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a.prefix[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  const int m = lo;
  const int* S = a.st + 4 * m;
  const int kp = S[2];
  if (kp <= 0) return;
  const int n = a.ns[m];
  update_tile<T, TM, KC>(a.A[m], a.lda[m], a.ws + a.wsoff[m], n, S[1], kp, n, b - a.prefix[m],
                         Us, Ws);
}

__global__ void k_batch_fill(int* p, int count, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

// A_m(i, j) = A_m(j, i) = scale * u(seed + m, i, j) (SplitMix64 uniform in
// [-1, 1)); A_m(i, i) = 0 when zero_diag_stride > 0 and i % stride == 0.
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

template <typename T>
__global__ void k_batched_gen(int count, const int* __restrict__ ns, unsigned long long seed,
                              double scale, int zstride, T* const* A,
                              const int64_t* __restrict__ lda) {
  for (int m = blockIdx.x; m < count; m += gridDim.x) {
    const int n = ns[m];
    T* Am = A[m];
    const int64_t ld = lda[m];
    const unsigned long long base = splitmix64(seed + static_cast<unsigned long long>(m));
    for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += blockDim.x) {
      const int i = static_cast<int>(e / n), j = static_cast<int>(e % n);
      const int lo = min(i, j), hi = max(i, j);
      const unsigned long long key =
          splitmix64(base ^ ((static_cast<unsigned long long>(lo) << 32) |
                             static_cast<unsigned long long>(hi)));
      double u = 2.0 * (static_cast<double>(key >> 11) * 0x1.0p-53) - 1.0;
      if (i == j && zstride > 0 && i % zstride == 0) u = 0.0;
      Am[(int64_t)i * ld + j] = static_cast<T>(scale * u);
    }
  }
}

// ---- host side -----------------------------------------------------------

template <typename T, int NT>
static int batched_grid_nt(Handle* h, int count) {
  int per_cu = 0, cus = 0;
  const int lds = batch_lds_bytes(NT);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batched_sytrf<T, NT>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void*>(&k_batched_sytrf<T, NT>), NT, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) !=
          hipSuccess ||
      cus < 1)
    cus = 1;
  return std::max(1, std::min(count, per_cu * cus));
}

template <typename T>
static int resident_grid(Handle* h, int count) {
  int per_cu = 0, cus = 0;
  const void* kern = reinterpret_cast<const void*>(&k_batched_lds<T>);
  (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                            kBatchResidentLdsBytes);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, kBatchResidentLdsBytes) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) !=
          hipSuccess ||
      cus < 1)
    cus = 1;
  return std::max(1, std::min(count, per_cu * cus));
}

template <typename T>
static int sym_grid(Handle* h, int count, int lds) {
  int per_cu = 0, cus = 0;
  const void* kern = reinterpret_cast<const void*>(&k_batched_sym<T>);
  (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) !=
          hipSuccess ||
      cus < 1)
    cus = 1;
  return std::max(1, std::min(count, per_cu * cus));
}

static int batched_grid(Handle* h, int count, int elem_bytes, int nt) {
  if (elem_bytes == 8)
    return nt == 64 ? batched_grid_nt<double, 64>(h, count) : batched_grid_nt<double, 256>(h, count);
  return nt == 64 ? batched_grid_nt<float, 64>(h, count) : batched_grid_nt<float, 256>(h, count);
}

// Lockstep schedule: per matrix a W slice and 4 state ints; per panel step
// the prefix sums of the matrices' trailing-tile counts (a bound from the
// nominal panel end: a deferred 2x2 ends a panel at most one column early).
// Dynamic LDS of k_step_panel<T, true> that keeps as many workgroups on a
// CU as its registers allow (160 KB of LDS per CU shared among them, less
// the kernel's static LDS); the kernel's limit is raised to it once.
static int step_panel_lds_cap(int elem_bytes) {
  auto cap_of = [](const void* k) {
    hipFuncAttributes fa;
    size_t stat = 16 * 1024;
    if (hipFuncGetAttributes(&fa, k) == hipSuccess)
      stat = fa.sharedSizeBytes;
    else
      (void)hipGetLastError();
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0) != hipSuccess ||
        per_cu < 1) {
      (void)hipGetLastError();
      per_cu = 4;
    }
    const size_t share = 160 * 1024 / static_cast<size_t>(per_cu);
    const int cap =
        share > stat + 1024 ? static_cast<int>((share - stat - 1024) / 1024 * 1024) : 0;
    // past 64 KB a launch needs the raised limit
    if (cap > 64 * 1024 &&
        hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, cap) != hipSuccess) {
      (void)hipGetLastError();
      return 64 * 1024;
    }
    return cap;
  };
  static const int cap64 = cap_of(reinterpret_cast<const void*>(&k_step_panel<double, true>));
  static const int cap32 = cap_of(reinterpret_cast<const void*>(&k_step_panel<float, true>));
  return elem_bytes == 8 ? cap64 : cap32;
}

static int prepare_lockstep(Handle* h, int count, const int* h_n, int elem_bytes,
                            const BatchWidths& wt) {
  BatchPlan& p = h->batch;
  int steps = 0;
  std::vector<int64_t> wsoff(count);
  int64_t wel = 0;
  for (int i = 0; i < count; ++i) {
    const int n = h_n[i], bw = wt.w[batch_class(n)];
    if (n > 0) steps = std::max(steps, (n + bw - 1) / bw);
    wsoff[i] = wel;
    wel += ((int64_t)(bw + 2) * n + 63) / 64 * 64;
  }
  const int64_t sstride = count + 1;
  std::vector<int> prefix((size_t)steps * sstride, 0);
  for (int P = 0; P < steps; ++P) {
    int* pr = prefix.data() + (size_t)P * sstride;
    int acc = 0;
    for (int i = 0; i < count; ++i) {
      pr[i] = acc;
      const int n = h_n[i], bw = wt.w[batch_class(n)];
      if (n > 0 && P < (n + bw - 1) / bw) {
        const int mb = std::max(0, n - std::min((P + 1) * bw, n) + 1);
        const int T = (mb + 63) / 64;
        acc += T * (T + 1) / 2;
      }
    }
    pr[count] = acc;
  }
  // buf: n[count], state[4 count], prefix[steps * (count + 1)],
  // order[count], wsoff (int64)
  const size_t ints = count + 4 * (size_t)count + prefix.size() + count + 64;
  const size_t bytes = ints * sizeof(int) + count * sizeof(int64_t) + 256;
  int rc = h->reserve_buf(&p.buf, &p.buf_bytes, bytes, "batch schedule");
  if (rc != BBK_OK) return rc;
  rc = h->reserve_buf(&p.ws, &p.ws_bytes, std::max<size_t>(1, wel) * elem_bytes, "batch workspace");
  if (rc != BBK_OK) return rc;
  std::vector<char> host(bytes, 0);
  int* hi = reinterpret_cast<int*>(host.data());
  std::copy(h_n, h_n + count, hi);
  std::copy(prefix.begin(), prefix.end(), hi + 5 * (size_t)count);
  // panel workgroups take the matrices largest first: a step lasts until
  // its slowest resident workgroups finish, and a batch larger than the
  // resident grid balances best when the long panels start first
  int* ord = hi + 5 * (size_t)count + prefix.size();
  std::iota(ord, ord + count, 0);
  std::stable_sort(ord, ord + count, [&](int x, int y) { return h_n[x] > h_n[y]; });
  const size_t off64 = ((ints * sizeof(int)) + 255) / 256 * 256;
  std::memcpy(host.data() + off64, wsoff.data(), count * sizeof(int64_t));
  hipError_t e = hipMemcpyAsync(p.buf, host.data(), bytes, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return h->hip_fail(e, "upload batch schedule");
  p.count = count;
  p.elem = elem_bytes;
  p.sizes.assign(h_n, h_n + count);
  p.widths = wt;
  p.lockstep = true;
  p.steps = steps;
  p.step_stride = sstride;
  p.ws_elems = wel;
  p.step_tiles.resize(steps);
  for (int P = 0; P < steps; ++P) p.step_tiles[P] = prefix[(size_t)P * sstride + count];
  p.grid = count;
  p.nt = 256;
  p.resident = false;
  p.sym = false;
  // the panel's pivot and candidate rows in LDS when both fit
  // kStepRowLdsBytes; beyond them, as far as the workgroups its registers
  // allow per CU still fit, copies of the first W rows
  int nmax = 0;
  for (int i = 0; i < count; ++i) nmax = std::max(nmax, h_n[i]);
  const size_t row_lds = 2 * static_cast<size_t>(nmax) * elem_bytes;
  p.lds_dyn = 0;
  if (nmax > 0 && row_lds <= kStepRowLdsBytes) {
    const size_t want = row_lds + static_cast<size_t>(kBatchMaxWidth) * nmax * elem_bytes;
    const size_t cap = static_cast<size_t>(step_panel_lds_cap(elem_bytes));
    p.lds_dyn = static_cast<int>(std::max(row_lds, std::min(want, cap)));
  }
  return BBK_OK;
}

// Sort largest first, upload sizes and queue order, size the global W
// slices for matrices whose panel rows do not fit LDS.
int batched_prepare(Handle* h, int count, const int* h_n, int elem_bytes) {
  if (count < 0) return -2;
  if (count > 0 && !h_n) return -3;
  if (elem_bytes != 4 && elem_bytes != 8) return -4;
  for (int i = 0; i < count; ++i)
    if (h_n[i] < 0) return -3;
  BatchPlan& p = h->batch;
  const BatchWidths& wt = h->batch_widths[elem_bytes == 8 ? 0 : 1];
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return h_n[a] > h_n[b]; });
  int nmax = 0;
  for (int i = 0; i < count; ++i) nmax = std::max(nmax, h_n[i]);
  if (nmax > kBatchLockstepN) return prepare_lockstep(h, count, h_n, elem_bytes, wt);
  if (nmax <= kBatchResidentN && wt.nt[batch_class(nmax)] == kBatchTeamSym) {
    // symmetric-workspace kernel: LDS for the widest panel of the batch's
    // classes, one n x n slice of S per workgroup
    int wmax = 1;
    for (int c = 0; c <= batch_class(std::max(nmax, 1)); ++c) wmax = std::max(wmax, wt.w[c]);
    if (wmax > kSymMaxWidth)
      return h->fail(BBK_ERR_UNSUPPORTED,
                     "batched: the symmetric-workspace mode takes panel widths up to 32");
    const int lds = sym_lds_bytes(std::max(nmax, 1), wmax, elem_bytes);
    int grid = elem_bytes == 8 ? sym_grid<double>(h, std::max(count, 1), lds)
                               : sym_grid<float>(h, std::max(count, 1), lds);
    if (h->batch_grid_limit > 0) grid = std::min(grid, h->batch_grid_limit);
    const int64_t stride = ((int64_t)std::max(nmax, 1) * std::max(nmax, 1) + 63) / 64 * 64;
    const size_t ints = 2 * static_cast<size_t>(count) + 64;
    int rc = h->reserve_buf(&p.buf, &p.buf_bytes, ints * sizeof(int), "batch schedule");
    if (rc != BBK_OK) return rc;
    rc = h->reserve_buf(&p.ws, &p.ws_bytes, static_cast<size_t>(stride) * grid * elem_bytes,
                        "batch workspace");
    if (rc != BBK_OK) return rc;
    std::vector<int> host(ints, 0);
    std::copy(order.begin(), order.end(), host.begin());
    std::copy(h_n, h_n + count, host.begin() + count);
    hipError_t e = hipMemcpyAsync(p.buf, host.data(), ints * sizeof(int), hipMemcpyHostToDevice,
                                  h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return h->hip_fail(e, "upload batch schedule");
    p.count = count;
    p.elem = elem_bytes;
    p.sizes.assign(h_n, h_n + count);
    p.widths = wt;
    p.ws_stride = stride;
    p.grid = grid;
    p.nt = 64;
    p.lockstep = false;
    p.resident = false;
    p.sym = true;
    p.lds_dyn = lds;
    return BBK_OK;
  }
  if (nmax <= kBatchResidentN && wt.nt[batch_class(nmax)] == 0) {
    int grid = elem_bytes == 8 ? resident_grid<double>(h, std::max(count, 1))
                               : resident_grid<float>(h, std::max(count, 1));
    if (h->batch_grid_limit > 0) grid = std::min(grid, h->batch_grid_limit);
    const size_t ints = 2 * static_cast<size_t>(count) + 64;
    int rc = h->reserve_buf(&p.buf, &p.buf_bytes, ints * sizeof(int), "batch schedule");
    if (rc != BBK_OK) return rc;
    std::vector<int> host(ints, 0);
    std::copy(order.begin(), order.end(), host.begin());
    std::copy(h_n, h_n + count, host.begin() + count);
    hipError_t e = hipMemcpyAsync(p.buf, host.data(), ints * sizeof(int), hipMemcpyHostToDevice,
                                  h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return h->hip_fail(e, "upload batch schedule");
    p.count = count;
    p.elem = elem_bytes;
    p.sizes.assign(h_n, h_n + count);
    p.widths = wt;
    p.ws_stride = 0;
    p.grid = grid;
    p.nt = 256;
    p.lockstep = false;
    p.resident = true;
    p.sym = false;
    return BBK_OK;
  }
  const int nt = wt.nt[batch_class(nmax)] == 64 ? 64 : 256;  // 0 (resident) above 256: four waves
  const int lds_elems = batch_lds_bytes(nt) / elem_bytes;
  int64_t stride = 0;
  for (int i = 0; i < count; ++i) {
    const int n = h_n[i];
    const int c = batch_class(n);
    const int64_t need = (int64_t)(wt.w[c] + 2) * n;
    if (need > lds_elems) stride = std::max(stride, need);
  }
  stride = (stride + 63) / 64 * 64;
  int grid = batched_grid(h, std::max(count, 1), elem_bytes, nt);
  if (h->batch_grid_limit > 0) grid = std::min(grid, h->batch_grid_limit);
  const size_t ints = 2 * static_cast<size_t>(count) + 64;
  int rc = h->reserve_buf(&p.buf, &p.buf_bytes, ints * sizeof(int), "batch schedule");
  if (rc != BBK_OK) return rc;
  const size_t wsb = static_cast<size_t>(stride) * grid * elem_bytes;
  if (wsb > 0) {
    rc = h->reserve_buf(&p.ws, &p.ws_bytes, wsb, "batch workspace");
    if (rc != BBK_OK) return rc;
  }
  std::vector<int> host(ints, 0);
  std::copy(order.begin(), order.end(), host.begin());
  std::copy(h_n, h_n + count, host.begin() + count);
  hipError_t e = hipMemcpyAsync(p.buf, host.data(), ints * sizeof(int), hipMemcpyHostToDevice,
                                h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return h->hip_fail(e, "upload batch schedule");
  p.count = count;
  p.elem = elem_bytes;
  p.sizes.assign(h_n, h_n + count);
  p.widths = wt;
  p.ws_stride = stride;
  p.grid = grid;
  p.nt = nt;
  p.lockstep = false;
  p.resident = false;
  p.sym = false;
  return BBK_OK;
}

static bool plan_matches(const BatchPlan& p, int count, const int* h_n, int elem_bytes,
                         const BatchWidths& wt) {
  if (p.count != count || p.elem != elem_bytes) return false;
  for (int c = 0; c < kBatchClasses; ++c)
    if (p.widths.w[c] != wt.w[c] || p.widths.nt[c] != wt.nt[c] || p.widths.ts[c] != wt.ts[c])
      return false;
  return count == 0 || std::equal(h_n, h_n + count, p.sizes.begin());
}

// Queue reset (fill kernel: graph-capture safe) + the persistent launch.
template <typename T>
static int lockstep_launch(Handle* h, const BatchPlan& p, T* const* A, const int64_t* lda,
                           int* const* perm, int* const* piv) {
  int* base = reinterpret_cast<int*>(p.buf);
  const size_t ints =
      p.count + 4 * (size_t)p.count + (size_t)p.steps * p.step_stride + p.count + 64;
  StepArgs<T> a;
  a.count = p.count;
  a.ns = base;
  a.A = A;
  a.lda = lda;
  a.perm = perm;
  a.piv = piv;
  a.ws = reinterpret_cast<T*>(p.ws);
  a.wsoff = reinterpret_cast<const int64_t*>(reinterpret_cast<char*>(p.buf) +
                                             ((ints * sizeof(int)) + 255) / 256 * 256);
  a.st = base + p.count;
  a.widths = p.widths;
  const int* prefix_all = base + 5 * (size_t)p.count;
  a.order = prefix_all + (size_t)p.steps * p.step_stride;
  // the pivot and candidate rows (and W row copies) in LDS: p.lds_dyn > 0
  a.lds_dyn = p.lds_dyn;
  for (int P = 0; P < p.steps; ++P) {
    a.step = P;
    a.prefix = prefix_all + (size_t)P * p.step_stride;
    if (p.lds_dyn > 0)
      hipLaunchKernelGGL((k_step_panel<T, true>), dim3(p.count), dim3(256), p.lds_dyn,
                         h->stream, a);
    else
      hipLaunchKernelGGL((k_step_panel<T, false>), dim3(p.count), dim3(256), 0, h->stream, a);
    const int tiles = p.step_tiles[P];
    if (tiles > 0)
      hipLaunchKernelGGL(k_step_update<T>, dim3(tiles), dim3(256), 0, h->stream, a);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return h->hip_fail(e, "lockstep batched launch");
  return BBK_OK;
}

template <typename T>
int batched_launch(Handle* h, const BatchPlan& p, T* const* A, const int64_t* lda,
                   int* const* perm, int* const* piv) {
  if (p.count == 0) return BBK_OK;
  if (p.lockstep) return lockstep_launch<T>(h, p, A, lda, perm, piv);
  int* base = reinterpret_cast<int*>(p.buf);
  BatchArgs<T> a;
  a.count = p.count;
  a.order = base;
  a.ns = base + p.count;
  a.A = A;
  a.lda = lda;
  a.perm = perm;
  a.piv = piv;
  a.queue = base + 2 * p.count;
  a.ws = reinterpret_cast<T*>(p.ws);
  a.ws_stride = p.ws_stride;
  a.lds_elems = batch_lds_bytes(p.nt) / static_cast<int>(sizeof(T));
  a.lds_bytes = kBatchResidentLdsBytes;
  a.widths = p.widths;
  hipLaunchKernelGGL(k_batch_fill, dim3(1), dim3(64), 0, h->stream, a.queue, 1, 0);
  if (p.sym)
    hipLaunchKernelGGL(k_batched_sym<T>, dim3(p.grid), dim3(64), p.lds_dyn, h->stream, a);
  else if (p.resident)
    hipLaunchKernelGGL(k_batched_lds<T>, dim3(p.grid), dim3(256), kBatchResidentLdsBytes,
                       h->stream, a);
  else if (p.nt == 64)
    hipLaunchKernelGGL((k_batched_sytrf<T, 64>), dim3(p.grid), dim3(64), batch_lds_bytes(64),
                       h->stream, a);
  else
    hipLaunchKernelGGL((k_batched_sytrf<T, 256>), dim3(p.grid), dim3(256), batch_lds_bytes(256),
                       h->stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return h->hip_fail(e, "batched factorization launch");
  return BBK_OK;
}

template <typename T>
int batched_sytrf(Handle* h, int count, const int* h_n, T* const* A, const int64_t* lda,
                  int* const* perm, int* const* piv) {
  if (count < 0) return -2;
  if (count > 0 && (!h_n)) return -3;
  if (count > 0 && !A) return -4;
  if (count > 0 && !lda) return -5;
  if (count > 0 && !perm) return -6;
  if (count > 0 && !piv) return -7;
  const int eb = sizeof(T);
  if (!plan_matches(h->batch, count, h_n, eb, h->batch_widths[eb == 8 ? 0 : 1])) {
    int rc = batched_prepare(h, count, h_n, eb);
    if (rc != BBK_OK) return rc;
  }
  return batched_launch<T>(h, h->batch, A, lda, perm, piv);
}

template int batched_sytrf<double>(Handle*, int, const int*, double* const*, const int64_t*,
                                   int* const*, int* const*);
template int batched_sytrf<float>(Handle*, int, const int*, float* const*, const int64_t*,
                                  int* const*, int* const*);

template <typename T>
int batched_gen(Handle* h, int count, const int* d_n, uint64_t seed, double scale,
                int zstride, T* const* A, const int64_t* lda) {
  if (count < 0) return -2;
  if (count == 0) return BBK_OK;
  if (!d_n) return -3;
  if (!A) return -7;
  if (!lda) return -8;
  const int grid = std::min(count, 4096);
  hipLaunchKernelGGL(k_batched_gen<T>, dim3(grid), dim3(256), 0, h->stream, count, d_n,
                     static_cast<unsigned long long>(seed), scale, zstride, A, lda);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return h->hip_fail(e, "batched generator launch");
  return BBK_OK;
}

// ---- autotuner -----------------------------------------------------------
// For every size class up to max_n: a synthetic batch with sizes spread
// over the class (in the resident range, from 16 up to the class bound;
// uniform [-1, 1] entries, zero diagonal on every 7th
// index to force 2x2 pivots), each candidate (panel width, team size) timed
// with HIP events (best of two), the fastest kept in the handle's table.
template <typename T>
int batched_tune(Handle* h, int max_n, int* widths_out) {
  static const int cand[] = {8, 16, 24, 32, 48, 64};
  const int eb = sizeof(T);
  BatchWidths& tab = h->batch_widths[eb == 8 ? 0 : 1];
  int lo = 1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return h->fail(BBK_ERR_HIP, "event create");
  int rc = BBK_OK;
  for (int c = 0; c < kBatchClasses && rc == BBK_OK; ++c) {
    const int hi = std::min(batch_class_bound(c), std::max(max_n, 1));
    if (lo > hi) break;
    // A batch runs with the mode of its largest matrix's class, so a class
    // of the resident range (n <= kBatchResidentN) is tuned on batches whose
    // largest class it is: sizes from 16 (or its own lower bound, if
    // smaller) up to its bound, the smaller ones under their classes' tuned
    // widths and tiles.  Enough matrices to keep every CU's teams busy for
    // small classes (concurrency decides the team size there), about
    // 2.5e10 element operations per trial batch for the large ones.
    const int lo_t = hi <= kBatchResidentN ? std::min(lo, 16) : lo;
    const double mean_n = 0.5 * (lo_t + hi);
    int count = static_cast<int>(std::min(3072.0, std::max(64.0, 2.5e10 / (mean_n * mean_n * mean_n))));
    std::vector<int> ns(count);
    for (int i = 0; i < count; ++i)
      ns[i] = lo_t + static_cast<int>((static_cast<int64_t>(i) * 7919) % (hi - lo_t + 1));
    std::vector<int64_t> lds(count);
    size_t elems = 0;
    std::vector<size_t> offs(count);
    for (int i = 0; i < count; ++i) {
      lds[i] = (ns[i] + 15) / 16 * 16;
      offs[i] = elems;
      elems += static_cast<size_t>(lds[i]) * ns[i];
    }
    void* mem = nullptr;
    const size_t ptr_bytes = static_cast<size_t>(count) * (4 * sizeof(void*) + sizeof(int64_t) + sizeof(int));
    const size_t bytes = 2 * elems * eb + 2 * static_cast<size_t>(count) * hi * sizeof(int) + ptr_bytes + 4096;
    if (hipMalloc(&mem, bytes) != hipSuccess) {
      rc = h->fail(BBK_ERR_ALLOC, "tuner buffers");
      break;
    }
    char* b = reinterpret_cast<char*>(mem);
    T* src = reinterpret_cast<T*>(b);
    T* work = src + elems;
    int* ip = reinterpret_cast<int*>(work + elems);
    std::vector<void*> hA(count), hS(count), hP(count), hV(count);
    for (int i = 0; i < count; ++i) {
      hS[i] = src + offs[i];
      hA[i] = work + offs[i];
      hP[i] = ip + static_cast<size_t>(i) * hi;
      hV[i] = ip + static_cast<size_t>(count) * hi + static_cast<size_t>(i) * hi;
    }
    char* pb = reinterpret_cast<char*>(ip + 2 * static_cast<size_t>(count) * hi);
    pb = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(pb) + 255) / 256 * 256);
    void** dA = reinterpret_cast<void**>(pb);
    void** dS = dA + count;
    void** dP = dS + count;
    void** dV = dP + count;
    int64_t* dL = reinterpret_cast<int64_t*>(dV + count);
    int* dN = reinterpret_cast<int*>(dL + count);
    hipError_t e = hipMemcpy(dA, hA.data(), count * sizeof(void*), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dS, hS.data(), count * sizeof(void*), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dP, hP.data(), count * sizeof(void*), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dV, hV.data(), count * sizeof(void*), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dL, lds.data(), count * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dN, ns.data(), count * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      hipFree(mem);
      rc = h->hip_fail(e, "tuner upload");
      break;
    }
    rc = batched_gen<T>(h, count, dN, 0x7475u + c, 1.0, 7, reinterpret_cast<T* const*>(dS), dL);
    float best_ms = 1e30f;
    int best_w = tab.w[c], best_nt = tab.nt[c];
    // classes n <= kBatchResidentN: panel width x mode (one-wavefront team,
    // four-wavefront team, resident in LDS, symmetric workspace) x trailing
    // update (SIMT 4 x 4, MFMA 16 x 16, MFMA 32 x 32; the resident and
    // symmetric-workspace modes are MFMA only); larger classes: panel width
    // x team size, SIMT update
    static const int rcand[] = {8, 16, 24, 32};
    static const int modes[] = {64, 256, 0, kBatchTeamSym};
    static const int tiles[] = {0, 16, 32};
    const bool res_class = hi <= kBatchResidentN;
    const int nw_small = static_cast<int>(sizeof(rcand) / sizeof(rcand[0]));
    const int ntrials = res_class ? nw_small * 4 * 3
                                  : 2 * static_cast<int>(sizeof(cand) / sizeof(cand[0]));
    int best_ts = tab.ts[c];
    for (int trial = 0; trial < ntrials; ++trial) {
      const int w = res_class ? rcand[trial / 12] : cand[trial >> 1];
      const int nt = res_class ? modes[(trial / 3) % 4] : ((trial & 1) ? 256 : 64);
      const int ts = res_class ? tiles[trial % 3] : 0;
      // the resident and symmetric-workspace kernels have MFMA updates only
      if (res_class && (nt == 0 || nt == kBatchTeamSym) && ts == 0) continue;
      if (rc != BBK_OK) break;
      if (w > kBatchMaxWidth) continue;
      BatchWidths saved = tab;
      tab.w[c] = w;
      tab.nt[c] = nt;
      tab.ts[c] = ts;
      float ms_min = 1e30f;
      rc = batched_prepare(h, count, ns.data(), eb);
      for (int rep = 0; rep < 2 && rc == BBK_OK; ++rep) {
        e = hipMemcpyAsync(work, src, elems * eb, hipMemcpyDeviceToDevice, h->stream);
        if (e != hipSuccess) { rc = h->hip_fail(e, "tuner copy"); break; }
        hipEventRecord(e0, h->stream);
        rc = batched_launch<T>(h, h->batch, reinterpret_cast<T* const*>(dA), dL,
                               reinterpret_cast<int* const*>(dP), reinterpret_cast<int* const*>(dV));
        hipEventRecord(e1, h->stream);
        e = hipEventSynchronize(e1);
        if (e != hipSuccess) { rc = h->hip_fail(e, "tuner run"); break; }
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        ms_min = std::min(ms_min, ms);
      }
      tab = saved;
      if (ms_min < best_ms) {
        best_ms = ms_min;
        best_w = w;
        best_nt = nt;
        best_ts = ts;
      }
    }
    tab.ts[c] = best_ts;
    tab.w[c] = best_w;
    tab.nt[c] = best_nt;
    hipFree(mem);
    lo = hi + 1;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  h->batch.count = -1;  // force a fresh plan with the tuned table
  if (widths_out)
    for (int c = 0; c < kBatchClasses; ++c) widths_out[c] = tab.w[c];
  return rc;
}

// A private copy of the handle's current plan for a graph: its own
// schedule buffer (copied from the handle's) and its own W slices, so
// later batched calls, re-plans or tuning on the handle cannot change what
// a captured graph reads.  Freed by release_plan.
static int clone_plan(Handle* h, BatchPlan* out) {
  const BatchPlan& p = h->batch;
  *out = p;
  out->buf = nullptr;
  out->ws = nullptr;
  hipError_t e = hipSuccess;
  if (p.buf_bytes > 0) {
    e = hipMalloc(&out->buf, p.buf_bytes);
    if (e == hipSuccess)
      e = hipMemcpyAsync(out->buf, p.buf, p.buf_bytes, hipMemcpyDeviceToDevice, h->stream);
  }
  if (e == hipSuccess && p.ws_bytes > 0) e = hipMalloc(&out->ws, p.ws_bytes);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    release_plan(out);
    return h->hip_fail(e, "graph plan copy");
  }
  return BBK_OK;
}

void release_plan(BatchPlan* p) {
  if (p->buf) (void)hipFree(p->buf);
  if (p->ws) (void)hipFree(p->ws);
  p->buf = nullptr;
  p->ws = nullptr;
  p->buf_bytes = p->ws_bytes = 0;
}

// Capture one batched factorization (queue reset + persistent kernel, or
// the lockstep launches) on the handle's stream into a graph that reads
// only the private plan `gp` (filled here; the caller keeps it alive as
// long as the graph and releases it after).
template <typename T>
int batched_graph(Handle* h, int count, const int* h_n, T* const* A, const int64_t* lda,
                  int* const* perm, int* const* piv, hipGraph_t* g, hipGraphExec_t* x,
                  BatchPlan* gp) {
  const int eb = sizeof(T);
  if (count < 0) return -2;
  if (count > 0 && (!h_n || !A || !lda || !perm || !piv)) return -3;
  if (!plan_matches(h->batch, count, h_n, eb, h->batch_widths[eb == 8 ? 0 : 1])) {
    int rc = batched_prepare(h, count, h_n, eb);
    if (rc != BBK_OK) return rc;
  }
  int rc = clone_plan(h, gp);
  if (rc != BBK_OK) return rc;
  hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    release_plan(gp);
    return h->hip_fail(e, "begin capture");
  }
  rc = batched_launch<T>(h, *gp, A, lda, perm, piv);
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(h->stream, &graph);
  if (rc != BBK_OK || e != hipSuccess) {
    if (graph) hipGraphDestroy(graph);
    release_plan(gp);
    return rc != BBK_OK ? rc : h->hip_fail(e, "end capture");
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(graph);
    release_plan(gp);
    return h->hip_fail(e, "graph instantiate");
  }
  *g = graph;
  *x = exec;
  return BBK_OK;
}

}  // namespace bbk

// A graph owns its copy of the batch plan (schedule and W slices).
struct bbk_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
  hipStream_t stream;
  bbk::BatchPlan plan;
};

using bbk::Handle;

extern "C" {

int bbk_dsytrf_batched(bbk_handle_t handle, int count, const int* h_n, double* const* d_A,
                       const int64_t* d_lda, int* const* d_perm, int* const* d_piv) {
  if (!handle) return -1;
  return bbk::batched_sytrf<double>(reinterpret_cast<Handle*>(handle), count, h_n, d_A, d_lda,
                                    d_perm, d_piv);
}

int bbk_ssytrf_batched(bbk_handle_t handle, int count, const int* h_n, float* const* d_A,
                       const int64_t* d_lda, int* const* d_perm, int* const* d_piv) {
  if (!handle) return -1;
  return bbk::batched_sytrf<float>(reinterpret_cast<Handle*>(handle), count, h_n, d_A, d_lda,
                                   d_perm, d_piv);
}

int bbk_batched_prepare(bbk_handle_t handle, int count, const int* h_n, int elem_bytes) {
  if (!handle) return -1;
  return bbk::batched_prepare(reinterpret_cast<Handle*>(handle), count, h_n, elem_bytes);
}

int bbk_batched_get_widths(bbk_handle_t handle, int elem_bytes, int* widths) {
  if (!handle) return -1;
  if (elem_bytes != 4 && elem_bytes != 8) return -2;
  if (!widths) return -3;
  const bbk::BatchWidths& t =
      reinterpret_cast<Handle*>(handle)->batch_widths[elem_bytes == 8 ? 0 : 1];
  for (int c = 0; c < bbk::kBatchClasses; ++c) widths[c] = t.w[c];
  return BBK_OK;
}

int bbk_batched_set_widths(bbk_handle_t handle, int elem_bytes, const int* widths) {
  if (!handle) return -1;
  if (elem_bytes != 4 && elem_bytes != 8) return -2;
  if (!widths) return -3;
  for (int c = 0; c < bbk::kBatchClasses; ++c)
    if (widths[c] < 1 || widths[c] > bbk::kBatchMaxWidth) return -3;
  Handle* h = reinterpret_cast<Handle*>(handle);
  bbk::BatchWidths& t = h->batch_widths[elem_bytes == 8 ? 0 : 1];
  for (int c = 0; c < bbk::kBatchClasses; ++c) t.w[c] = widths[c];
  h->batch.count = -1;
  return BBK_OK;
}

int bbk_batched_get_team_sizes(bbk_handle_t handle, int elem_bytes, int* team_sizes) {
  if (!handle) return -1;
  if (elem_bytes != 4 && elem_bytes != 8) return -2;
  if (!team_sizes) return -3;
  const bbk::BatchWidths& t =
      reinterpret_cast<Handle*>(handle)->batch_widths[elem_bytes == 8 ? 0 : 1];
  for (int c = 0; c < bbk::kBatchClasses; ++c) team_sizes[c] = t.nt[c];
  return BBK_OK;
}

int bbk_batched_set_team_sizes(bbk_handle_t handle, int elem_bytes, const int* team_sizes) {
  if (!handle) return -1;
  if (elem_bytes != 4 && elem_bytes != 8) return -2;
  if (!team_sizes) return -3;
  for (int c = 0; c < bbk::kBatchClasses; ++c)
    if (team_sizes[c] != 64 && team_sizes[c] != 256 &&
        !((team_sizes[c] == 0 || team_sizes[c] == bbk::kBatchTeamSym) &&
          bbk::batch_class_bound(c) <= bbk::kBatchResidentN))
      return -3;
  Handle* h = reinterpret_cast<Handle*>(handle);
  bbk::BatchWidths& t = h->batch_widths[elem_bytes == 8 ? 0 : 1];
  for (int c = 0; c < bbk::kBatchClasses; ++c) t.nt[c] = team_sizes[c];
  h->batch.count = -1;
  return BBK_OK;
}

int bbk_batched_get_tiles(bbk_handle_t handle, int elem_bytes, int* tiles) {
  if (!handle) return -1;
  if (elem_bytes != 4 && elem_bytes != 8) return -2;
  if (!tiles) return -3;
  const bbk::BatchWidths& t =
      reinterpret_cast<Handle*>(handle)->batch_widths[elem_bytes == 8 ? 0 : 1];
  for (int c = 0; c < bbk::kBatchClasses; ++c) tiles[c] = t.ts[c];
  return BBK_OK;
}

int bbk_batched_set_tiles(bbk_handle_t handle, int elem_bytes, const int* tiles) {
  if (!handle) return -1;
  if (elem_bytes != 4 && elem_bytes != 8) return -2;
  if (!tiles) return -3;
  for (int c = 0; c < bbk::kBatchClasses; ++c)
    if (tiles[c] != 0 && tiles[c] != 16 && tiles[c] != 32) return -3;
  Handle* h = reinterpret_cast<Handle*>(handle);
  bbk::BatchWidths& t = h->batch_widths[elem_bytes == 8 ? 0 : 1];
  for (int c = 0; c < bbk::kBatchClasses; ++c) t.ts[c] = tiles[c];
  h->batch.count = -1;
  return BBK_OK;
}

int bbk_batched_set_grid_limit(bbk_handle_t handle, int max_workgroups) {
  if (!handle) return -1;
  if (max_workgroups < 0) return -2;
  Handle* h = reinterpret_cast<Handle*>(handle);
  h->batch_grid_limit = max_workgroups;
  h->batch.count = -1;  // the next batched call plans again
  return BBK_OK;
}

int bbk_batched_tune(bbk_handle_t handle, int elem_bytes, int max_n, int* widths_out) {
  if (!handle) return -1;
  if (elem_bytes != 4 && elem_bytes != 8) return -2;
  if (max_n < 1) return -3;
  Handle* h = reinterpret_cast<Handle*>(handle);
  return elem_bytes == 8 ? bbk::batched_tune<double>(h, max_n, widths_out)
                         : bbk::batched_tune<float>(h, max_n, widths_out);
}

int bbk_graph_create_dsytrf_batched(bbk_handle_t handle, int count, const int* h_n,
                                    double* const* d_A, const int64_t* d_lda,
                                    int* const* d_perm, int* const* d_piv,
                                    bbk_graph_t* graph) {
  if (!handle) return -1;
  if (!graph) return -8;
  Handle* h = reinterpret_cast<Handle*>(handle);
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  bbk::BatchPlan plan;
  int rc = bbk::batched_graph<double>(h, count, h_n, d_A, d_lda, d_perm, d_piv, &g, &x, &plan);
  if (rc != BBK_OK) return rc;
  *graph = new bbk_graph{g, x, h->stream, plan};
  return BBK_OK;
}

int bbk_graph_create_ssytrf_batched(bbk_handle_t handle, int count, const int* h_n,
                                    float* const* d_A, const int64_t* d_lda,
                                    int* const* d_perm, int* const* d_piv,
                                    bbk_graph_t* graph) {
  if (!handle) return -1;
  if (!graph) return -8;
  Handle* h = reinterpret_cast<Handle*>(handle);
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  bbk::BatchPlan plan;
  int rc = bbk::batched_graph<float>(h, count, h_n, d_A, d_lda, d_perm, d_piv, &g, &x, &plan);
  if (rc != BBK_OK) return rc;
  *graph = new bbk_graph{g, x, h->stream, plan};
  return BBK_OK;
}

int bbk_graph_launch(bbk_graph_t graph) {
  if (!graph) return -1;
  return hipGraphLaunch(graph->exec, graph->stream) == hipSuccess ? BBK_OK : BBK_ERR_HIP;
}

int bbk_graph_destroy(bbk_graph_t graph) {
  if (!graph) return -1;
  // a replay still queued must finish before its plan's buffers go
  (void)hipStreamSynchronize(graph->stream);
  if (graph->exec) hipGraphExecDestroy(graph->exec);
  if (graph->graph) hipGraphDestroy(graph->graph);
  bbk::release_plan(&graph->plan);
  delete graph;
  return BBK_OK;
}

int bbk_dgen_synth_batched(bbk_handle_t handle, int count, const int* d_n, uint64_t seed,
                           double scale, int zero_diag_stride, double* const* d_A,
                           const int64_t* d_lda) {
  if (!handle) return -1;
  return bbk::batched_gen<double>(reinterpret_cast<Handle*>(handle), count, d_n, seed, scale,
                                  zero_diag_stride, d_A, d_lda);
}

int bbk_sgen_synth_batched(bbk_handle_t handle, int count, const int* d_n, uint64_t seed,
                           double scale, int zero_diag_stride, float* const* d_A,
                           const int64_t* d_lda) {
  if (!handle) return -1;
  return bbk::batched_gen<float>(reinterpret_cast<Handle*>(handle), count, d_n, seed, scale,
                                 zero_diag_stride, d_A, d_lda);
}

}  // extern "C"

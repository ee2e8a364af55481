// bbk_batched_refine.hip -- batched mixed-precision solve: binary32
// factorization of every matrix, then binary64 iterative refinement of
// A X = B per matrix, all matrices in persistent launches.
//
// Refinement kernel: one workgroup (four wavefronts) per matrix, taken from
// a work queue ordered largest first.  Per matrix it runs
//   x = solve32(b);  loop { r = b - A x (binary64);  eta_q = ||r_q||_inf /
//   (||A||_inf ||x_q||_inf + ||b_q||_inf);  stop when max_q eta_q <= tol or
//   after max_iter corrections;  x += solve32(r) }
// with the right-hand sides (up to four) kept in LDS: binary32 copies for
// the triangular sweeps, binary64 accumulators for A x.  Both sweeps are
// right-looking over 64-row blocks: one wavefront solves a diagonal block
// serially (readlane broadcast), the whole workgroup applies the block to
// the rest, one row of the factor per thread.  The 2x2
// D blocks use LAPACK dsytrs's scaled formula in binary64.  A x reads each
// stored row of the upper triangle once: the row part is a wavefront dot
// product, the mirrored column part is added into LDS accumulators with
// LDS atomics (so its rounding order may vary between runs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "../../include/bbk_ldlt.h"
#include "bbk_batched.hpp"
#include "bbk_handle.hpp"

namespace bbk {

constexpr int kRT = 256;          // threads per matrix team
constexpr int kRB = 64;           // sweep block
constexpr int kRMaxRhs = 4;

struct RefineArgs {
  int count;
  int nrhs;
  const int* order;
  const int* ns;
  const double* const* A;
  const int64_t* lda;
  const float* const* F;
  const int64_t* ldf;
  const int* const* perm;
  const int* const* piv;
  const double* const* B;
  const int64_t* ldb;
  double* const* X;
  const int64_t* ldx;
  int* queue;
  double tol;
  int max_iter;
  int* iters;
  double* berr;
};

extern __shared__ __attribute__((aligned(16))) unsigned char bbk_refine_lds[];

__device__ __forceinline__ float rl(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// z (n x NR, LDS, row-major) := L^-T D^-1 L^-1 z with L, D from the
// upper-stored binary32 factor (L(j, c) = F(c, j) for c < j, except
// L(k+1, k) = 0 inside a 2x2 block, whose F(k, k+1) holds D).
// Inlined into its one call site: as a call it spilled 740 bytes per lane
// to scratch (1024-matrix workload 134.8 -> 131.5 ms inlined).
template <int NR>
__device__ __forceinline__ void solve32(int n, const float* __restrict__ F, int64_t ldf,
                        const int* __restrict__ piv, float* z, float* part) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int nb = (n + kRB - 1) / kRB;
  // forward: z := L^-1 z, right-looking over 64-column blocks
  for (int J = 0; J < nb; ++J) {
    const int j0 = J * kRB, jn = min(kRB, n - j0);
    if (w == 0) {
      const int j = j0 + lane;
      const bool ok = lane < jn;
      float Lc[kRB];
#pragma unroll
      for (int ii = 0; ii < kRB; ++ii)
        Lc[ii] = (ok && ii < lane) ? F[(int64_t)(j0 + ii) * ldf + j] : 0.0f;
      if (ok && lane > 0 && piv[j - 1] == 2) {
#pragma unroll
        for (int ii = 0; ii < kRB; ++ii)
          if (ii == lane - 1) Lc[ii] = 0.0f;
      }
      float v[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q) v[q] = ok ? z[j * NR + q] : 0.0f;
#pragma unroll
      for (int ii = 0; ii < kRB - 1; ++ii) {
        if (ii < jn - 1) {
#pragma unroll
          for (int q = 0; q < NR; ++q) v[q] -= Lc[ii] * rl(v[q], ii);
        }
      }
      if (ok)
#pragma unroll
        for (int q = 0; q < NR; ++q) z[j * NR + q] = v[q];
    }
    __syncthreads();
    // rows below the block; row j0 + jn couples to column j0 + jn - 1 only
    // through L, which is 0 when a 2x2 block spans the two
    const bool cut = j0 + jn < n && piv[j0 + jn - 1] == 2;
    for (int i = j0 + jn + t; i < n; i += kRT) {
      float acc[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q) acc[q] = 0.0f;
      const int jend = (cut && i == j0 + jn) ? jn - 1 : jn;
      // all of the row's loads in flight before the multiply-adds
      float lv[kRB];
#pragma unroll
      for (int jj = 0; jj < kRB; ++jj) lv[jj] = jj < jend ? F[(int64_t)(j0 + jj) * ldf + i] : 0.0f;
#pragma unroll
      for (int jj = 0; jj < kRB; ++jj) {
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] += lv[jj] * z[(j0 + jj) * NR + q];
      }
#pragma unroll
      for (int q = 0; q < NR; ++q) z[i * NR + q] -= acc[q];
    }
    __syncthreads();
  }
  // D^-1 (binary64 arithmetic; 2x2 blocks with LAPACK dsytrs's formula)
  for (int i = t; i < n; i += kRT) {
    const int pv = piv[i];
    if (pv == 2 || pv == -2) {
      const int k = pv == 2 ? i : i - 1;
      const double a = F[(int64_t)k * ldf + k];
      const double b = F[(int64_t)k * ldf + k + 1];
      const double c = F[(int64_t)(k + 1) * ldf + k + 1];
      const double akm1 = a / b, ak = c / b;
      const double denom = akm1 * ak - 1.0;
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        const double bkm1 = (double)z[k * NR + q] / b;
        const double bk = (double)z[(k + 1) * NR + q] / b;
        part[i * NR + q] =
            (float)(pv == 2 ? (ak * bkm1 - bk) / denom : (akm1 * bk - bkm1) / denom);
      }
    } else {
      const double d = F[(int64_t)i * ldf + i];
#pragma unroll
      for (int q = 0; q < NR; ++q) part[i * NR + q] = (float)((double)z[i * NR + q] / d);
    }
  }
  __syncthreads();
  for (int i = t; i < n * NR; i += kRT) z[i] = part[i];
  __syncthreads();
  // backward: z := L^-T z, right-looking over 64-row blocks from the last:
  // wavefront 0 solves the diagonal block, then every row j above it takes
  // z_j -= sum_{i in block} L(i, j) z_i, L(i, j) = F(j, i): one row of F
  // per thread, 64 independent loads
  for (int I = nb - 1; I >= 0; --I) {
    const int i0 = I * kRB, in = min(kRB, n - i0);
    if (w == 0) {
      const int i = i0 + lane;
      const bool ok = lane < in;
      float Lr[kRB];
#pragma unroll
      for (int jj = 0; jj < kRB; ++jj)
        Lr[jj] = (ok && jj > lane && jj < in) ? F[(int64_t)i * ldf + i0 + jj] : 0.0f;
      if (ok && lane + 1 < in && piv[i] == 2) {
#pragma unroll
        for (int jj = 0; jj < kRB; ++jj)
          if (jj == lane + 1) Lr[jj] = 0.0f;
      }
      float v[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q) v[q] = ok ? z[i * NR + q] : 0.0f;
#pragma unroll
      for (int jj = kRB - 1; jj > 0; --jj) {
        if (jj < in) {
#pragma unroll
          for (int q = 0; q < NR; ++q) v[q] -= Lr[jj] * rl(v[q], jj);
        }
      }
      if (ok)
#pragma unroll
        for (int q = 0; q < NR; ++q) z[i * NR + q] = v[q];
    }
    __syncthreads();
    // row i0 - 1 couples to column i0 only through L, which is 0 when a
    // 2x2 block spans the two
    const bool cut = i0 > 0 && piv[i0 - 1] == 2;
    for (int j = t; j < i0; j += kRT) {
      const int ifirst = (cut && j == i0 - 1) ? 1 : 0;
      float acc[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q) acc[q] = 0.0f;
      // the row's 64 entries in two halves of 32 loads in flight
#pragma unroll
      for (int h = 0; h < kRB; h += kRB / 2) {
        float lv[kRB / 2];
#pragma unroll
        for (int u = 0; u < kRB / 2; ++u) {
          const int ii = h + u;
          lv[u] = (ii >= ifirst && ii < in) ? F[(int64_t)j * ldf + i0 + ii] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kRB / 2; ++u) {
#pragma unroll
          for (int q = 0; q < NR; ++q) acc[q] += lv[u] * z[(i0 + h + u) * NR + q];
        }
      }
#pragma unroll
      for (int q = 0; q < NR; ++q) z[j * NR + q] -= acc[q];
    }
    __syncthreads();
  }
}

// y[i NR + q] += sum_j A(i, j) x_j[q] for the symmetric A stored in the
// upper triangle (ABS: |A(i, j)| into y[i NR], x ignored).  Wavefront w
// takes row blocks of RB rows; per 64-column chunk each lane (one column)
// accumulates the block's row-part products in registers and its column
// part for the chunk, so LDS atomics happen once per RB rows per column and
// once per row at the end of the block (not per element).
template <int NR, bool ABS>
__device__ __forceinline__ void team_symv(int n, const double* __restrict__ A, int64_t lda,
                                          const double* __restrict__ X, int64_t ldx,
                                          double* y) {
  constexpr int QN = ABS ? 1 : NR;
  constexpr int RB = QN <= 2 ? 16 : 8;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int jb = w * RB; jb < n; jb += (kRT / 64) * RB) {
    double xj[RB][QN], racc[RB][QN];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int q = 0; q < QN; ++q) {
        racc[r][q] = 0.0;
        xj[r][q] = ABS ? 1.0 : (jb + r < n ? X[(int64_t)(jb + r) * ldx + q] : 0.0);
      }
    // the chunk's RB entries are loaded one chunk ahead
    double an[RB];
    {
      const int i = jb + lane;
#pragma unroll
      for (int r = 0; r < RB; ++r) an[r] = (i < n && i >= jb + r) ? A[(int64_t)(jb + r) * lda + i] : 0.0;
    }
    for (int ib = jb; ib < n; ib += 64) {
      const int i = ib + lane;
      const bool vi = i < n;
      double ac[RB];
#pragma unroll
      for (int r = 0; r < RB; ++r) ac[r] = an[r];
      {
        const int i2 = i + 64;
#pragma unroll
        for (int r = 0; r < RB; ++r)
          an[r] = (i2 < n && i2 >= jb + r) ? A[(int64_t)(jb + r) * lda + i2] : 0.0;
      }
      double xi[QN], cacc[QN];
#pragma unroll
      for (int q = 0; q < QN; ++q) {
        xi[q] = ABS ? 1.0 : (vi ? X[(int64_t)i * ldx + q] : 0.0);
        cacc[q] = 0.0;
      }
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int j = jb + r;
        double a = ac[r];  // (vi && i >= j) ? A(j, i) : 0; i >= j implies j < n
        if (ABS) a = fabs(a);
#pragma unroll
        for (int q = 0; q < QN; ++q) racc[r][q] += a * xi[q];
        if (i > j) {
#pragma unroll
          for (int q = 0; q < QN; ++q) cacc[q] += a * xj[r][q];
        }
      }
      if (vi) {
#pragma unroll
        for (int q = 0; q < QN; ++q) atomicAdd(&y[i * NR + q], cacc[q]);
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int q = 0; q < QN; ++q) {
        const double sr = wave_sum(racc[r][q]);
        if (lane == 0 && jb + r < n) atomicAdd(&y[(jb + r) * NR + q], sr);
      }
  }
}

template <int NR>
__device__ __forceinline__ double team_max_abs(double v, double* red) {
  v = wave_max_nan_ignoring(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double m = red[0];
  for (int k = 1; k < kRT / 64; ++k) m = (m < red[k]) ? red[k] : m;
  __syncthreads();
  return m;
}

template <int NR>
__global__ __launch_bounds__(kRT) void k_batched_refine(RefineArgs a) {
  __shared__ int sq;
  __shared__ double red[kRT / 64];
  __shared__ double anorm_s;
  for (;;) {
    if (threadIdx.x == 0) sq = atomicAdd(a.queue, 1);
    __syncthreads();
    const int qpos = sq;
    __syncthreads();
    if (qpos >= a.count) break;
    const int m = a.order[qpos];
    const int n = a.ns[m];
    if (n <= 0) {
      // an empty system is solved exactly: no correction, zero error
      if (threadIdx.x == 0) {
        a.iters[m] = 0;
        a.berr[m] = 0.0;
      }
      continue;
    }
    const double* A = a.A[m];
    const int64_t lda = a.lda[m];
    const float* F = a.F[m];
    const int64_t ldf = a.ldf[m];
    const int* perm = a.perm[m];
    const int* piv = a.piv[m];
    const double* Bm = a.B[m];
    const int64_t ldb = a.ldb[m];
    double* Xm = a.X[m];
    const int64_t ldx = a.ldx[m];
    double* y = reinterpret_cast<double*>(bbk_refine_lds);        // n x NR
    float* z = reinterpret_cast<float*>(y + (int64_t)n * NR);      // n x NR
    float* part = z + (int64_t)n * NR;                             // n x NR
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;

    // x0 = solve32(b), then corrections x += solve32(r); one call site
    // keeps the sweep code inlined once
    int it = 0;
    double eta = 0.0;
    bool first = true;
    for (;;) {
      if (!first) {
        // y = A x, binary64; ||A||_inf on the first pass (y[., 0] reused)
        const bool want_norm = it == 0;
        if (want_norm) {
          for (int i = t; i < n; i += kRT) y[i * NR] = 0.0;
          __syncthreads();
          team_symv<NR, true>(n, A, lda, nullptr, 0, y);
          __syncthreads();
          double mx = 0.0;
          for (int i = t; i < n; i += kRT) mx = (mx < y[i * NR]) ? y[i * NR] : mx;
          mx = team_max_abs<NR>(mx, red);
          if (t == 0) anorm_s = mx;
          __syncthreads();
        }
        for (int i = t; i < n * NR; i += kRT) y[i] = 0.0;
        __syncthreads();
        team_symv<NR, false>(n, A, lda, Xm, ldx, y);
        __syncthreads();
        // r = b - y (kept in y), eta
        double rmax[NR], xmax[NR], bmax[NR];
  #pragma unroll
        for (int q = 0; q < NR; ++q) rmax[q] = xmax[q] = bmax[q] = 0.0;
        for (int i = t; i < n; i += kRT) {
  #pragma unroll
          for (int q = 0; q < NR; ++q) {
            const double b = Bm[(int64_t)i * ldb + q];
            const double r = b - y[i * NR + q];
            y[i * NR + q] = r;
            const double ar = fabs(r), ax = fabs(Xm[(int64_t)i * ldx + q]), ab = fabs(b);
            rmax[q] = (rmax[q] < ar) ? ar : rmax[q];
            xmax[q] = (xmax[q] < ax) ? ax : xmax[q];
            bmax[q] = (bmax[q] < ab) ? ab : bmax[q];
            if (!(r == r)) rmax[q] = __builtin_huge_val();
          }
        }
        eta = 0.0;
  #pragma unroll
        for (int q = 0; q < NR; ++q) {
          const double R = team_max_abs<NR>(rmax[q], red);
          const double Xn = team_max_abs<NR>(xmax[q], red);
          const double Bn = team_max_abs<NR>(bmax[q], red);
          const double den = anorm_s * Xn + Bn;
          const double e = den > 0.0 ? R / den : (R > 0.0 ? __builtin_huge_val() : 0.0);
          eta = (eta < e || !(e == e)) ? e : eta;
        }
        if (eta <= a.tol || it >= a.max_iter || !(eta == eta)) break;
      }
      // x (+)= P solve32(P^T rhs), rhs = b first, then r (kept in y)
      for (int i = t; i < n; i += kRT) {
        const int pi = perm[i];
#pragma unroll
        for (int q = 0; q < NR; ++q)
          z[i * NR + q] = first ? (float)Bm[(int64_t)pi * ldb + q] : (float)y[pi * NR + q];
      }
      __syncthreads();
      solve32<NR>(n, F, ldf, piv, z, part);
      for (int i = t; i < n; i += kRT) {
        const int pi = perm[i];
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          double* xp = &Xm[(int64_t)pi * ldx + q];
          *xp = first ? (double)z[i * NR + q] : *xp + (double)z[i * NR + q];
        }
      }
      __syncthreads();
      if (!first) ++it;
      first = false;
    }
    if (t == 0) {
      a.iters[m] = it;
      a.berr[m] = eta;
    }
    __syncthreads();
  }
}

// Upper triangle binary64 -> binary32.  Workgroup g of the grid takes
// matrix g % count and every (gridDim.x / count)-th group of 4 rows of it
// (a grid of a multiple of count), one wavefront per row, four 64-column
// chunks of loads in flight per lane.
__global__ __launch_bounds__(256) void k_batched_cast(int count, const int* __restrict__ ns,
                                                      const double* const* A,
                                                      const int64_t* __restrict__ lda,
                                                      float* const* S,
                                                      const int64_t* __restrict__ lds) {
  const int per = gridDim.x / count;  // workgroups per matrix
  const int m = blockIdx.x % count, part = blockIdx.x / count;
  const int n = ns[m];
  const double* Am = A[m];
  float* Sm = S[m];
  const int64_t la = lda[m], ls = lds[m];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = 4 * part + w; i < n; i += 4 * per) {
    const double* Ar = Am + (int64_t)i * la;
    float* Sr = Sm + (int64_t)i * ls;
    int j = i + lane;
    for (; j + 192 < n; j += 256) {
      const double a0 = Ar[j], a1 = Ar[j + 64], a2 = Ar[j + 128], a3 = Ar[j + 192];
      Sr[j] = (float)a0;
      Sr[j + 64] = (float)a1;
      Sr[j + 128] = (float)a2;
      Sr[j + 192] = (float)a3;
    }
    for (; j < n; j += 64) Sr[j] = (float)Ar[j];
  }
}

__global__ void k_refine_fill(int* p, int v) { *p = v; }

static size_t align256(size_t v) { return (v + 255) / 256 * 256; }

int batched_sysv_mixed(Handle* h, int count, const int* h_n, int nrhs,
                       const double* const* A, const int64_t* lda, const double* const* B,
                       const int64_t* ldb, double* const* X, const int64_t* ldx, double tol,
                       int max_iter, int* iters, double* berr) {
  if (count < 0) return -2;
  if (count == 0) return BBK_OK;
  if (!h_n) return -3;
  if (nrhs < 1 || nrhs > kRMaxRhs) return -4;
  if (!A) return -5;
  if (!lda) return -6;
  if (!B) return -7;
  if (!ldb) return -8;
  if (!X) return -9;
  if (!ldx) return -10;
  if (max_iter < 0) return -12;
  if (!iters) return -13;
  if (!berr) return -14;
  int nmax = 0;
  for (int i = 0; i < count; ++i) {
    if (h_n[i] < 0) return -3;
    nmax = std::max(nmax, h_n[i]);
  }
  const size_t lds_bytes = static_cast<size_t>(nmax) * nrhs * (sizeof(double) + 2 * sizeof(float));
  if (static_cast<size_t>(nmax) * nrhs > 9600)
    return h->fail(BBK_ERR_UNSUPPORTED, "batched refinement: n * nrhs too large for LDS");
  // binary32 factor copies + perm/piv, pointer arrays
  std::vector<int64_t> ldf(count);
  std::vector<size_t> foff(count), ioff(count);
  size_t fel = 0, iel = 0;
  for (int i = 0; i < count; ++i) {
    ldf[i] = std::max<int64_t>(1, (h_n[i] + 63) / 64 * 64);
    foff[i] = fel;
    fel += static_cast<size_t>(ldf[i]) * h_n[i];
    ioff[i] = iel;
    iel += static_cast<size_t>(h_n[i]);
  }
  const size_t fbytes = align256(fel * sizeof(float));
  const size_t ibytes = align256(2 * iel * sizeof(int) + 64);
  const size_t pbytes = align256(static_cast<size_t>(count) * (3 * sizeof(void*) + sizeof(int64_t)) +
                                 (2 * static_cast<size_t>(count) + 16) * sizeof(int));
  int rc = h->reserve_mixed(fbytes + ibytes + pbytes);
  if (rc != BBK_OK) return rc;
  char* base = reinterpret_cast<char*>(h->mixed_ws);
  float* fbase = reinterpret_cast<float*>(base);
  int* ibase = reinterpret_cast<int*>(base + fbytes);
  char* pb = base + fbytes + ibytes;
  std::vector<char> host(pbytes, 0);
  void** hF = reinterpret_cast<void**>(host.data());
  void** hP = hF + count;
  void** hV = hP + count;
  int64_t* hL = reinterpret_cast<int64_t*>(hV + count);
  int* hOrder = reinterpret_cast<int*>(hL + count);
  int* hN = hOrder + count;
  int* hQueue = hN + count;
  for (int i = 0; i < count; ++i) {
    hF[i] = fbase + foff[i];
    hP[i] = ibase + ioff[i];
    hV[i] = ibase + iel + ioff[i];
    hL[i] = ldf[i];
  }
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return h_n[x] > h_n[y]; });
  std::copy(order.begin(), order.end(), hOrder);
  std::copy(h_n, h_n + count, hN);
  *hQueue = 0;
  hipError_t e = hipMemcpyAsync(pb, host.data(), pbytes, hipMemcpyHostToDevice, h->stream);
  if (e != hipSuccess) return h->hip_fail(e, "upload refinement schedule");
  auto dptr = [&](void* hp) { return pb + (reinterpret_cast<char*>(hp) - host.data()); };
  float* const* dF = reinterpret_cast<float* const*>(dptr(hF));
  int* const* dP = reinterpret_cast<int* const*>(dptr(hP));
  int* const* dV = reinterpret_cast<int* const*>(dptr(hV));
  const int64_t* dL = reinterpret_cast<const int64_t*>(dptr(hL));
  const int* dOrder = reinterpret_cast<const int*>(dptr(hOrder));
  const int* dN = reinterpret_cast<const int*>(dptr(hN));
  int* dQueue = reinterpret_cast<int*>(dptr(hQueue));
  // 8 workgroups per matrix (4 rows each per pass), at least 2048 in all
  const int cast_per = count >= (1 << 27) ? 1 : std::max(8, 2048 / count);
  hipLaunchKernelGGL(k_batched_cast, dim3(count * cast_per), dim3(256), 0, h->stream, count, dN,
                     A, lda, dF, dL);
  rc = batched_sytrf<float>(h, count, h_n, dF, dL, dP, dV);
  if (rc != BBK_OK) return rc;
  RefineArgs ra;
  ra.count = count;
  ra.nrhs = nrhs;
  ra.order = dOrder;
  ra.ns = dN;
  ra.A = A;
  ra.lda = lda;
  ra.F = dF;
  ra.ldf = dL;
  ra.perm = dP;
  ra.piv = dV;
  ra.B = B;
  ra.ldb = ldb;
  ra.X = X;
  ra.ldx = ldx;
  ra.queue = dQueue;
  ra.tol = tol;
  ra.max_iter = max_iter;
  ra.iters = iters;
  ra.berr = berr;
  int cus = 1;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device);
  hipLaunchKernelGGL(k_refine_fill, dim3(1), dim3(1), 0, h->stream, dQueue, 0);
  const void* kern = nullptr;
  switch (nrhs) {
    case 1: kern = reinterpret_cast<const void*>(&k_batched_refine<1>); break;
    case 2: kern = reinterpret_cast<const void*>(&k_batched_refine<2>); break;
    case 3: kern = reinterpret_cast<const void*>(&k_batched_refine<3>); break;
    default: kern = reinterpret_cast<const void*>(&k_batched_refine<4>); break;
  }
  (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds_bytes));
  int per_cu = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kRT, lds_bytes) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  int grid = std::max(1, std::min(count, per_cu * cus));
  if (h->batch_grid_limit > 0) grid = std::min(grid, h->batch_grid_limit);
  switch (nrhs) {
    case 1:
      hipLaunchKernelGGL(k_batched_refine<1>, dim3(grid), dim3(kRT), lds_bytes, h->stream, ra);
      break;
    case 2:
      hipLaunchKernelGGL(k_batched_refine<2>, dim3(grid), dim3(kRT), lds_bytes, h->stream, ra);
      break;
    case 3:
      hipLaunchKernelGGL(k_batched_refine<3>, dim3(grid), dim3(kRT), lds_bytes, h->stream, ra);
      break;
    default:
      hipLaunchKernelGGL(k_batched_refine<4>, dim3(grid), dim3(kRT), lds_bytes, h->stream, ra);
      break;
  }
  e = hipGetLastError();
  if (e != hipSuccess) return h->hip_fail(e, "batched refinement launch");
  std::vector<double> hb(count);
  e = hipMemcpyAsync(hb.data(), berr, count * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return h->hip_fail(e, "batched refinement");
  for (double v : hb)
    if (!(v <= tol)) return BBK_ERR_NOT_CONVERGED;
  return BBK_OK;
}

}  // namespace bbk

extern "C" int bbk_dsysv_mixed_batched(bbk_handle_t handle, int count, const int* h_n, int nrhs,
                                       const double* const* d_A, const int64_t* d_lda,
                                       const double* const* d_B, const int64_t* d_ldb,
                                       double* const* d_X, const int64_t* d_ldx, double tol,
                                       int max_iter, int* d_iters, double* d_berr) {
  if (!handle) return -1;
  return bbk::batched_sysv_mixed(reinterpret_cast<bbk::Handle*>(handle), count, h_n, nrhs, d_A,
                                 d_lda, d_B, d_ldb, d_X, d_ldx, tol, max_iter, d_iters, d_berr);
}

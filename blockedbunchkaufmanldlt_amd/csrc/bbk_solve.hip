// bbk_solve.hip -- solves with the compact factors, the binary64 residual /
// backward error, and the mixed-precision refinement driver.
//
// Solve: x = P L^-T D^-1 L^-1 P^T b with (L D L^T)(i,j) = A(perm[i],perm[j]).
// L is stored transposed in the upper triangle (L(j,i) = F(i,j), i < j,
// except F(i,i+1) of a 2x2 pivot, which is D).  Both triangular sweeps
// run as one cooperative launch each: 128-row blocks, workgroups take
// blocks in dependency order and publish finished unknowns as tagged
// records (each half block in the substitution form, the whole block in
// the refinement's inverse form).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "../../include/bbk_ldlt.h"
#include "bbk_handle.hpp"
#include "bbk_kernels.hpp"

namespace bbk {

constexpr int kSolveBlock = 128;
constexpr int kSolveThreads = 256;
constexpr int kSolveMaxRhs = 4;  // right-hand sides per launch

__device__ __forceinline__ double readlane_d(double v, int lane) {
  int2 u = __builtin_bit_cast(int2, v);
  u.x = __builtin_amdgcn_readlane(u.x, lane);
  u.y = __builtin_amdgcn_readlane(u.y, lane);
  return __builtin_bit_cast(double, u);
}
__device__ __forceinline__ float readlane_f(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
template <typename T>
__device__ __forceinline__ T readlane_t(T v, int lane);
template <>
__device__ __forceinline__ double readlane_t<double>(double v, int lane) {
  return readlane_d(v, lane);
}
template <>
__device__ __forceinline__ float readlane_t<float>(float v, int lane) {
  return readlane_f(v, lane);
}

// Right-hand-side values cross workgroups as XRec records {tag, q, value}
// (one per row and right-hand side, index row * kSolveMaxRhs + q): a
// reader polls the record itself, so publishing needs no flag, fence or
// second round trip.  Tags are unique per launch.

// UB blocks at once: out[u] holds row row0[u] + (lane >> 2), rhs (lane & 3)
// of block u (0 for q >= NR or rows >= nlim).  All UB loads of a lane are
// in flight together and waited for in the same asm block.
template <int UB>
__device__ __forceinline__ void xload_multi(const XRec* const (&p)[UB], v4i (&d)[UB]);
template <>
__device__ __forceinline__ void xload_multi<2>(const XRec* const (&p)[2], v4i (&d)[2]) {
  asm volatile(
      "global_load_dwordx4 %0, %2, off sc1\n\t"
      "global_load_dwordx4 %1, %3, off sc1\n\t"
      "s_waitcnt vmcnt(0)"
      : "=&v"(d[0]), "=&v"(d[1])
      : "v"(p[0]), "v"(p[1])
      : "memory");
}
template <>
__device__ __forceinline__ void xload_multi<4>(const XRec* const (&p)[4], v4i (&d)[4]) {
  asm volatile(
      "global_load_dwordx4 %0, %4, off sc1\n\t"
      "global_load_dwordx4 %1, %5, off sc1\n\t"
      "global_load_dwordx4 %2, %6, off sc1\n\t"
      "global_load_dwordx4 %3, %7, off sc1\n\t"
      "s_waitcnt vmcnt(0)"
      : "=&v"(d[0]), "=&v"(d[1]), "=&v"(d[2]), "=&v"(d[3])
      : "v"(p[0]), "v"(p[1]), "v"(p[2]), "v"(p[3])
      : "memory");
}

template <int NR, int UB>
__device__ __forceinline__ void poll_rows_multi(const XRec* rec, const int (&row0)[UB], int nlim,
                                                int tag, double (&out)[UB], const WaitGuard& g) {
  const int lane = threadIdx.x & 63;
  const int q = lane & 3, rr = lane >> 2;
  const int qq = q < NR ? q : NR - 1;
  const XRec* p[UB];
  bool use[UB];
#pragma unroll
  for (int u = 0; u < UB; ++u) {
    const int row = row0[u] + rr;
    use[u] = q < NR && row < nlim;
    p[u] = rec + (int64_t)(row < nlim ? row : nlim - 1) * kSolveMaxRhs + qq;
    out[u] = 0.0;
  }
  // bounded (WaitGuard): past the deadline the missing values stay 0 and
  // the error is marked
  if (g.forced) {
    g.fail(kErrSweepWait);
    return;
  }
  for (unsigned polls = 1;; ++polls) {
    v4i d[UB];
    xload_multi<UB>(p, d);
    bool done = true;
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      if (!use[u]) continue;
      if (d[u].x == tag)
        out[u] = __longlong_as_double((static_cast<long long>(d[u].w) << 32) |
                                      static_cast<unsigned int>(d[u].z));
      else
        done = false;
    }
    if (__ballot(!done) == 0) break;
    if ((polls & 63u) == 0 && wall_clock64() > g.deadline) {
      g.fail(kErrSweepWait);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

__device__ __forceinline__ double bcast(double v, int lane) { return readlane_d(v, lane); }

// The diagonal block's unit L, strictly lower part packed by columns of
// the upper-triangle storage: Lp[c (c - 1) / 2 + r] = L(c, r) = F(b0 + r,
// b0 + c), r < c < bn, with the D entry F(k, k+1) of a 2x2 pivot read as 0
// (rows or columns >= bn read as 0).  The 128 x 128 tile is read as rows,
// 32 loads per thread in flight together, in two halves.
template <typename T>
__device__ __forceinline__ void stage_diag_block(T* Lp, const T* __restrict__ F, int64_t lda,
                                                 const int* __restrict__ piv, int b0, int bn,
                                                 int n) {
  constexpr int NB = kSolveBlock;
  constexpr int PER = NB * NB / kSolveThreads / 2;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    T v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + kSolveThreads * (u + PER * h);
      const int r = e / NB, c = e % NB;
      v[u] = F[(int64_t)min(b0 + r, n - 1) * lda + min(b0 + c, n - 1)];
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + kSolveThreads * (u + PER * h);
      const int r = e / NB, c = e % NB;
      if (r < c) {
        T x = c < bn ? v[u] : T(0);
        if (c == r + 1 && c < bn && piv[b0 + r] == 2) x = T(0);
        Lp[c * (c - 1) / 2 + r] = x;
      }
    }
  }
}

// D^-1 z for row i (binary64 arithmetic): a 1x1 pivot divides by d; row k or
// k+1 of a 2x2 pivot (k, k+1) uses LAPACK dsytrs's scaled formula.
template <typename T, int NR>
__device__ __forceinline__ void dsolve_row(const T* __restrict__ F, int64_t lda,
                                           const int* __restrict__ piv,
                                           const T* __restrict__ Z, int64_t ldz, int i,
                                           T (&wv)[NR]) {
  const int pv = piv[i];
  if (pv == 2 || pv == -2) {
    const int k = pv == 2 ? i : i - 1;
    const double a = F[(int64_t)k * lda + k];
    const double b = F[(int64_t)k * lda + k + 1];
    const double c = F[(int64_t)(k + 1) * lda + k + 1];
    const double akm1 = a / b, ak = c / b;
    const double denom = akm1 * ak - 1.0;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const double bkm1 = (double)Z[(int64_t)k * ldz + q] / b;
      const double bk = (double)Z[(int64_t)(k + 1) * ldz + q] / b;
      wv[q] = (T)(pv == 2 ? (ak * bkm1 - bk) / denom : (akm1 * bk - bkm1) / denom);
    }
  } else {
    const double d = F[(int64_t)i * lda + i];
#pragma unroll
    for (int q = 0; q < NR; ++q) wv[q] = (T)((double)Z[(int64_t)i * ldz + q] / d);
  }
}

// Inverse of each 128 x 128 unit lower diagonal block of L: Li[J] row-major,
// Li(r, c) = (L_JJ^-1)(r, c), zero above the diagonal and outside the
// matrix.  L_JJ is read like stage_diag_block (the D entry F(k, k+1) of a
// 2x2 pivot as 0); thread c forms column c by forward substitution, its
// column of the inverse kept in LDS (other threads never read it).
template <typename T>
__global__ __launch_bounds__(kSolveBlock) void k_diag_inv(int n, const T* __restrict__ F,
                                                          int64_t lda, const int* __restrict__ piv,
                                                          T* __restrict__ Li) {
  constexpr int NB = kSolveBlock;
  __shared__ T Ls[NB][NB];      // L(r, k), k < r
  __shared__ T Xs[NB][NB + 1];  // (L^-1)(i, c) at Xs[i][c]
  const int J = blockIdx.x;
  const int b0 = J * NB, bn = min(NB, n - b0);
  const int c = threadIdx.x;
  // row k of F holds L(., k): Ls[r][k] = F(b0 + k, b0 + r) for k < r
#pragma unroll 16
  for (int k = 0; k < NB; ++k) {
    T x = T(0);
    if (k < c && c < bn) {
      x = F[(int64_t)(b0 + k) * lda + b0 + c];
      if (c == k + 1 && piv[b0 + k] == 2) x = T(0);
    }
    Ls[c][k] = x;
  }
  __syncthreads();
  for (int i = 0; i < NB; ++i) {
    T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
    int k = 0;
    for (; k + 4 <= i; k += 4) {
      a0 += Ls[i][k] * Xs[k][c];
      a1 += Ls[i][k + 1] * Xs[k + 1][c];
      a2 += Ls[i][k + 2] * Xs[k + 2][c];
      a3 += Ls[i][k + 3] * Xs[k + 3][c];
    }
    for (; k < i; ++k) a0 += Ls[i][k] * Xs[k][c];
    const T acc = (a0 + a1) + (a2 + a3);
    Xs[i][c] = i < c ? T(0) : (i == c ? T(1) : -acc);
  }
  T* out = Li + (int64_t)J * NB * NB;
  for (int i = 0; i < NB; ++i) out[(int64_t)i * NB + c] = (i < bn && c < bn) ? Xs[i][c] : T(0);
}

// Next-block products of the binary32 refinement solves, both in one
// launch (blockIdx.y 0: forward, 1: backward; blockIdx.x = b < nblk - 1):
//
// forward, J = b + 1: G_J = Li_J L_{J,J-1}.  The forward sweep forms u_J =
// Li_J (b_J - sum_{I < J-1} L_JI z_I) before block J-1 is solved, then z_J
// = u_J - G_J z_{J-1}.  L_{J,J-1}(r, c) = F(j0 - 128 + c, j0 + r).
// Stored row-major at Li + (nblk + J) 128^2.
//
// backward, I = b: H_I = Li_I^T L_{I+1,I}^T.  The backward sweep forms u_I =
// Li_I^T (D^-1 z_I - sum_{J > I+1} L_JI^T v_J), then v_I = u_I - H_I v_{I+1}.
// L_{I+1,I}(c, k) = F(i0 + k, i0 + 128 + c).  Stored row-major at Li +
// (2 nblk + I) 128^2.
//
// The tiles read as 0 for rows past n and for the D entry of a 2x2 pivot
// across the block edge (F(j0 - 1, j0), F(i0 + 127, i0 + 128)).  Thread
// (c, h) forms column c of the product for rows 32h .. 32h+31; Li's block
// and the tile staged in LDS.
template <typename T>
__global__ __launch_bounds__(4 * kSolveBlock) void k_next_block(int n, const T* __restrict__ F,
                                                                int64_t lda,
                                                                const int* __restrict__ piv,
                                                                T* __restrict__ Li, int nblk) {
  constexpr int NB = kSolveBlock;
  __shared__ T Ls[NB][NB + 1];  // the diagonal block's inverse, Li(i, k)
  __shared__ T Lt[NB][NB + 1];  // the tile: F(p0 + k, q0 + c) at Lt[k][c]
  const bool fwd = blockIdx.y == 0;
  const int b = blockIdx.x;
  const int D = fwd ? b + 1 : b;       // block whose inverse is used
  const int p0 = b * NB, q0 = p0 + NB;  // tile rows (block b), columns (block b+1)
  const int qn = min(NB, n - q0);
  const int c = threadIdx.x & (NB - 1), h = threadIdx.x >> 7;
  const T* Lj = Li + (int64_t)D * NB * NB;
  const bool edge = piv[q0 - 1] == 2;
#pragma unroll 8
  for (int kk = 0; kk < NB / 4; ++kk) {
    const int k = 32 * h + kk;
    Ls[k][c] = Lj[(int64_t)k * NB + c];
    T x = c < qn ? F[(int64_t)(p0 + k) * lda + q0 + c] : T(0);
    if (k == NB - 1 && c == 0 && edge) x = T(0);
    Lt[k][c] = x;
  }
  __syncthreads();
  T* out = Li + (int64_t)((fwd ? nblk : 2 * nblk) + D) * NB * NB;
  for (int ii = 0; ii < NB / 4; ++ii) {
    const int i = 32 * h + ii;
    T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
    if (fwd) {
      // G(i, c) = sum_{k <= i} Li(i, k) L_{J,J-1}(k, c), L_{J,J-1}(k, c) =
      // F(p0 + c, q0 + k) = Lt[c][k]
      int k = 0;
      for (; k + 4 <= i + 1; k += 4) {
        a0 += Ls[i][k] * Lt[c][k];
        a1 += Ls[i][k + 1] * Lt[c][k + 1];
        a2 += Ls[i][k + 2] * Lt[c][k + 2];
        a3 += Ls[i][k + 3] * Lt[c][k + 3];
      }
      for (; k <= i; ++k) a0 += Ls[i][k] * Lt[c][k];
    } else {
      // H(i, c) = sum_{k >= i} Li(k, i) L_{I+1,I}(c, k), L_{I+1,I}(c, k) =
      // F(p0 + k, q0 + c) = Lt[k][c]
      int k = i;
      for (; k + 4 <= NB; k += 4) {
        a0 += Ls[k][i] * Lt[k][c];
        a1 += Ls[k + 1][i] * Lt[k + 1][c];
        a2 += Ls[k + 2][i] * Lt[k + 2][c];
        a3 += Ls[k + 3][i] * Lt[k + 3][c];
      }
      for (; k < NB; ++k) a0 += Ls[k][i] * Lt[k][c];
    }
    out[(int64_t)i * NB + c] = (a0 + a1) + (a2 + a3);
  }
}

// Strictly lower triangle of F := mirror of the strictly upper one (F(j, i)
// = F(i, j), i < j), through 64 x 64 LDS tiles so that reads and writes
// both run along rows.  Tile t of the upper triangle: (I, J), I <= J.
template <typename T>
__global__ __launch_bounds__(256) void k_mirror_upper(int n, T* __restrict__ F, int64_t lda) {
  constexpr int TB = 64;
  __shared__ T tile[TB][TB + 1];
  const int t = blockIdx.x;
  int J = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (J * (J + 1) / 2 > t) --J;
  while ((J + 1) * (J + 2) / 2 <= t) ++J;
  const int I = t - J * (J + 1) / 2;
  const int i0 = I * TB, j0 = J * TB;
  const int c = threadIdx.x & 63, r0 = threadIdx.x >> 6;
  for (int r = r0; r < TB; r += 4) {
    const int i = i0 + r, j = j0 + c;
    tile[r][c] = (i < n && j < n) ? F[(int64_t)i * lda + j] : T(0);
  }
  __syncthreads();
  // row j0 + r of the lower triangle, columns i0 + c: F(i0 + c, j0 + r)
  for (int r = r0; r < TB; r += 4) {
    const int j = j0 + r, i = i0 + c;
    if (j < n && i < j) F[(int64_t)j * lda + i] = tile[c][r];
  }
}

// Li (128 x 128 row-major, binary32) into LDS, rows of kInvLd elements
// (16-byte aligned): Lp[r kInvLd + c] = Li(r, c), or Li(c, r) when TRANS, so
// that both sweeps read rows of the staged matrix.
constexpr int kInvLd = kSolveBlock + 4;
template <typename T, bool TRANS>
__device__ __forceinline__ void stage_inverse_block(T* Lp, const T* __restrict__ Li) {
  constexpr int NB = kSolveBlock;
  constexpr int PER = NB * NB / kSolveThreads / 2;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    T v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) v[u] = Li[threadIdx.x + kSolveThreads * (u + PER * h)];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + kSolveThreads * (u + PER * h);
      const int r = e / NB, c = e % NB;
      Lp[TRANS ? c * kInvLd + r : r * kInvLd + c] = v[u];
    }
  }
}

// Half of a product with the staged inverse: thread (r = tid % 128, h = tid
// / 128) sums Lp[r][c] t[c] over c in [64 h, 64 h + 64), reading its row
// 16 bytes at a time (lanes on consecutive rows, 528 bytes apart: each
// group of 16 lanes covers the 64 banks once); t[c] is a broadcast read.
template <typename T, int NR>
__device__ __forceinline__ void inverse_product(const T* Lp, const T (*t)[NR], T (&sv)[NR]) {
  static_assert(sizeof(T) == 4, "binary32 inverse blocks");
  constexpr int NB = kSolveBlock;
  typedef float f4 __attribute__((ext_vector_type(4)));
  const int r = threadIdx.x & (NB - 1), c0 = (threadIdx.x >> 7) * 64;
  const f4* row = reinterpret_cast<const f4*>(Lp + r * kInvLd + c0);
  T a[2][NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) a[0][q] = a[1][q] = T(0);
#pragma unroll 4
  for (int c4 = 0; c4 < 16; ++c4) {
    const f4 l = row[c4];
    const int cc = c0 + 4 * c4;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      a[0][q] += l.x * t[cc][q];
      a[1][q] += l.y * t[cc + 1][q];
      a[0][q] += l.z * t[cc + 2][q];
      a[1][q] += l.w * t[cc + 3][q];
    }
  }
#pragma unroll
  for (int q = 0; q < NR; ++q) sv[q] = a[0][q] + a[1][q];
}

// Serial substitution over 64 unknowns, one per lane (v): step k takes
// unknown s = k (ASC) or 63 - k, broadcasts it with readlane and lanes with
// s < lane (ASC) or s > lane subtract L * v_s, L = Lp[idx(s)].  The L values
// come from LDS 16 at a time, the next 16 loaded ahead of the chain.
template <typename T, int NR, bool ASC, typename Idx>
__device__ __forceinline__ void serial64(T (&v)[NR], const T* Lp, int lane, Idx idx) {
  T cur[16], nxt[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) cur[u] = Lp[idx(ASC ? u : 63 - u)];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (c < 3) {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int k = 16 * (c + 1) + u;
        nxt[u] = Lp[idx(ASC ? k : 63 - k)];
      }
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int k = 16 * c + u;
      const int sidx = ASC ? k : 63 - k;
      const T l = (ASC ? sidx < lane : sidx > lane) ? cur[u] : T(0);
#pragma unroll
      for (int q = 0; q < NR; ++q) v[q] -= l * readlane_t<T>(v[q], sidx);
    }
    if (c < 3) {
#pragma unroll
      for (int u = 0; u < 16; ++u) cur[u] = nxt[u];
    }
  }
}

// Forward sweep: Z = L^-1 P^T B for NR (<= kSolveMaxRhs) columns.
// Workgroup takes 128-row blocks J = blockIdx.x, blockIdx.x + gridDim.x, ...
// (the diagonal block staged in LDS first).  Lane l holds columns j0 + l
// (A) and j0 + 64 + l (B); wavefront w accumulates rows 32w .. 32w+31 of
// every earlier block I < J, polling their records 16 rows at a time.  Then
// wavefront 0 solves the top 64 unknowns serially (readlane broadcast) and
// publishes them at once, so later blocks start on them; wavefront 1
// applies them to the bottom 64 and solves those.
//
// INV: the diagonal block's inverse Li (k_diag_inv) replaces the serial
// substitution, and G_J = Li_J L_{J,J-1} (k_next_block) the last
// block's contribution: u_J = Li_J (b_J - sum_{I < J-1} L_JI z_I) is formed
// before block J-1 is solved, then z_J = u_J - G_J z_{J-1}, so one product
// (thread (r, h) sums row r over columns 64h .. 64h+63) separates block
// J-1's records from block J's.
template <typename T, int NR, bool INV>
__global__ __launch_bounds__(kSolveThreads) void k_trsv_fwd(
    int n, const T* __restrict__ F, int64_t lda, const int* __restrict__ piv,
    const int* __restrict__ perm, const T* __restrict__ B, int64_t ldb,
    T* __restrict__ Z, int64_t ldz, XRec* __restrict__ rec, int tag, int nblk,
    const T* __restrict__ Li, const int* __restrict__ stop, unsigned* __restrict__ status,
    long long wait_ticks) {
  constexpr int NB = kSolveBlock, H = NB / 2;
  constexpr int NPK = INV ? NB * kInvLd : NB * (NB - 1) / 2;
  __shared__ __attribute__((aligned(16))) T Lp[NPK];  // packed L, or Li (rows of kInvLd)
  __shared__ T part[4][NB][NR];
  __shared__ T ztop[INV ? NB : H][NR];
  // INV: G_J = Li_J L_{J,J-1} (k_next_block, stored after the nblk
  // inverse blocks), staged when the workgroup starts: once block J-1's
  // records arrive, z_J = u_J - G_J z_{J-1} is one product away
  __shared__ __attribute__((aligned(16))) T Gs[INV ? NB * kInvLd : 4];
  if (stop && *stop) return;  // every workgroup reads the same flag
  const WaitGuard wg(status, wait_ticks);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int LAST = INV ? 1 : 0;  // blocks applied after the main loop
  for (int J = blockIdx.x; J < nblk; J += gridDim.x) {
    const int j0 = J * NB;
    const int jn = min(NB, n - j0);
    const int cA = j0 + lane, cB = j0 + H + lane;
    const bool okA = lane < jn, okB = H + lane < jn;
    if constexpr (INV) {
      stage_inverse_block<T, false>(Lp, Li + (int64_t)J * NB * NB);
      if (J > 0) stage_inverse_block<T, false>(Gs, Li + (int64_t)(nblk + J) * NB * NB);
    } else {
      stage_diag_block<T>(Lp, F, lda, piv, j0, jn, n);
    }
    T bj[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) bj[q] = T(0);
    if (w == 0 && okA) {
      const int pj = perm[cA];
#pragma unroll
      for (int q = 0; q < NR; ++q) bj[q] = B[(int64_t)pj * ldb + q];
    } else if (w == 1 && okB) {
      const int pj = perm[cB];
#pragma unroll
      for (int q = 0; q < NR; ++q) bj[q] = B[(int64_t)pj * ldb + q];
    }
    // 2x2 pivot across the block edge: L(j0, j0 - 1) = 0 (row j0 - 1 is row
    // 127 of block J - 1: wavefront 3, second chunk, last row; column j0 is
    // lane 0's column A)
    const bool edge2 = J > 0 && w == 3 && lane == 0 && piv[j0 - 1] == 2;
    T acc[2][NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[0][q] = acc[1][q] = T(0);
    // UB earlier blocks per step (2 UB chunks of 16 rows): F tiles and
    // records in flight together, one wait per step
    constexpr int UB = sizeof(T) == 8 ? 1 : 2;
    const int cAc = min(cA, n - 1), cBc = min(cB, n - 1);
    const int Jm = J - LAST;  // blocks I < Jm in the main loop
    for (int I = 0; I < Jm; I += UB) {
      T fa[2 * UB][16], fb[2 * UB][16];
      int row0[2 * UB];
#pragma unroll
      for (int k = 0; k < 2 * UB; ++k) {
        const int Iu = I + k / 2;
        const int rb = Iu * NB + 32 * w + 16 * (k & 1);
#pragma unroll
        for (int ii = 0; ii < 16; ++ii) {
          const int64_t ro = (int64_t)min(rb + ii, n - 1) * lda;
          fa[k][ii] = F[ro + cAc];
          fb[k][ii] = F[ro + cBc];
        }
        row0[k] = min(Iu, Jm - 1) * NB + 32 * w + 16 * (k & 1);
      }
#pragma unroll
      for (int k = 0; k < 2 * UB; ++k) {
        const bool live = I + k / 2 < Jm;
#pragma unroll
        for (int ii = 0; ii < 16; ++ii) {
          if (!(live && okA)) fa[k][ii] = T(0);
          if (!(live && okB)) fb[k][ii] = T(0);
        }
        if (edge2 && I + k / 2 == J - 1 && (k & 1)) fa[k][15] = T(0);
      }
      double zv[2 * UB];
      poll_rows_multi<NR, 2 * UB>(rec, row0, n, tag, zv, wg);
#pragma unroll
      for (int k = 0; k < 2 * UB; ++k)
#pragma unroll
        for (int ii = 0; ii < 16; ++ii) {
#pragma unroll
          for (int q = 0; q < NR; ++q) {
            const T z = static_cast<T>(bcast(zv[k], ii * 4 + q));
            acc[0][q] += fa[k][ii] * z;
            acc[1][q] += fb[k][ii] * z;
          }
        }
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      part[w][lane][q] = acc[0][q];
      part[w][H + lane][q] = acc[1][q];
    }
    __syncthreads();
    if constexpr (INV) {
      // t' into ztop (rows 0..127), then u = Li t'
      if (w < 2) {
        const int r = w * H + lane;
        const bool ok = w == 0 ? okA : okB;
#pragma unroll
        for (int q = 0; q < NR; ++q)
          ztop[r][q] = ok ? bj[q] - ((part[0][r][q] + part[1][r][q]) +
                                     (part[2][r][q] + part[3][r][q]))
                          : T(0);
      }
      __syncthreads();
      T sv[NR];
      inverse_product<T, NR>(Lp, ztop, sv);
      const int r = threadIdx.x & (NB - 1);
      if (threadIdx.x >= NB) {
#pragma unroll
        for (int q = 0; q < NR; ++q) part[0][r][q] = sv[q];
      }
      __syncthreads();
      T u[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q) u[q] = sv[q] + part[0][r][q];
      if (J > 0) {
        // block J-1's solution: wavefront w polls rows 32w .. 32w+31 into
        // ztop (free again), then u -= G_J z_{J-1}
        const int rb = (J - 1) * NB + 32 * w;
        const int row0[2] = {rb, rb + 16};
        double zv[2];
        poll_rows_multi<NR, 2>(rec, row0, n, tag, zv, wg);
        if ((lane & 3) < NR) {
#pragma unroll
          for (int k = 0; k < 2; ++k)
            ztop[32 * w + 16 * k + (lane >> 2)][lane & 3] = static_cast<T>(zv[k]);
        }
        __syncthreads();
        T gv[NR];
        inverse_product<T, NR>(Gs, ztop, gv);
        if (threadIdx.x >= NB) {
#pragma unroll
          for (int q = 0; q < NR; ++q) part[1][r][q] = gv[q];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NR; ++q) u[q] -= gv[q] + part[1][r][q];
      }
      if (threadIdx.x < NB && r < jn) {
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          xput(rec + (int64_t)(j0 + r) * kSolveMaxRhs + q, tag, q, static_cast<double>(u[q]));
          Z[(int64_t)(j0 + r) * ldz + q] = u[q];
        }
      }
      __syncthreads();  // Lp, Gs, part, ztop are reused by the next block
      continue;
    }
    if (w == 0) {
      T v[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q)
        v[q] = okA ? bj[q] - ((part[0][lane][q] + part[1][lane][q]) +
                              (part[2][lane][q] + part[3][lane][q]))
                   : T(0);
      const int cb = lane * (lane - 1) / 2;  // L(lane, r) at Lp[cb + r], r < lane
      serial64<T, NR, true>(v, Lp, lane, [&](int r) { return min(cb + r, NPK - 1); });
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        ztop[lane][q] = v[q];
        if (okA) {
          Z[(int64_t)cA * ldz + q] = v[q];
          xput(rec + (int64_t)cA * kSolveMaxRhs + q, tag, q, static_cast<double>(v[q]));
        }
      }
    }
    __syncthreads();
    if (w == 1) {
      T v[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q)
        v[q] = okB ? bj[q] - ((part[0][H + lane][q] + part[1][H + lane][q]) +
                              (part[2][H + lane][q] + part[3][H + lane][q]))
                   : T(0);
      const int c = H + lane;
      const int cb = c * (c - 1) / 2;  // L(c, r) at Lp[cb + r], r < c
      // the top unknowns applied in chunks of 8 (loads ahead of products)
#pragma unroll
      for (int i0 = 0; i0 < H; i0 += 8) {
        T lc[8], zc[8][NR];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          lc[u] = Lp[cb + i0 + u];
#pragma unroll
          for (int q = 0; q < NR; ++q) zc[u][q] = ztop[i0 + u][q];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int q = 0; q < NR; ++q) v[q] -= lc[u] * zc[u][q];
      }
      serial64<T, NR, true>(v, Lp, lane, [&](int r) { return min(cb + H + r, NPK - 1); });
      if (okB) {
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          Z[(int64_t)cB * ldz + q] = v[q];
          xput(rec + (int64_t)cB * kSolveMaxRhs + q, tag, q, static_cast<double>(v[q]));
        }
      }
    }
    __syncthreads();  // Lp, part, ztop are reused by the next block
  }
}

// Backward sweep with the D solve fused in front:
// V = L^-T D^-1 Z, X[perm[i]] = V[i].  Block I consumes blocks J > I; lane
// l holds rows i0 + l (A) and i0 + 64 + l (B), wavefront w covers columns
// 32w .. 32w+31 of each block J (16 at a time, contiguous per row).  The
// bottom 64 unknowns are solved first (wavefront 1) and published, then
// applied to the top 64, which wavefront 0 solves.
//
// INV (refinement solves): u_I = Li_I^T s' with Li the diagonal block's
// inverse (k_diag_inv), s' = D^-1 z_I - sum_{J > I+1} L_JI^T v_J, formed
// before block I+1 is solved; then v_I = u_I - H_I v_{I+1} with H_I =
// Li_I^T L_{I+1,I}^T (k_next_block).  Thread (r, h) sums row r of the
// staged Li^T or H_I over columns 64h .. 64h+63.  The strictly lower triangle of F then
// holds the mirror of the upper one (k_mirror_upper), and L_JI^T is read
// from it: element F(i, j), i in I, as F[j lda + i], consecutive lanes on
// consecutive i (coalesced, where the upper rows give every lane its own
// cache line).
template <typename T, int NR, bool INV>
__global__ __launch_bounds__(kSolveThreads) void k_trsv_bwd(
    int n, const T* __restrict__ F, int64_t lda, const int* __restrict__ piv,
    const int* __restrict__ perm, const T* __restrict__ Z, int64_t ldz,
    T* __restrict__ X, int64_t ldx, XRec* __restrict__ rec, int tag, int nblk,
    const T* __restrict__ Li, const int* __restrict__ stop, unsigned* __restrict__ status,
    long long wait_ticks) {
  constexpr int NB = kSolveBlock, H = NB / 2;
  constexpr int NPK = INV ? NB * kInvLd : NB * (NB - 1) / 2;
  __shared__ __attribute__((aligned(16))) T Lp[NPK];  // packed L, or Li (rows of kInvLd)
  __shared__ T part[4][NB][NR];
  __shared__ T vbot[INV ? NB : H][NR];
  // INV: H_I = Li_I^T L_{I+1,I}^T (k_next_block), staged when the
  // workgroup starts: once block I+1's records arrive, v_I = u_I - H_I
  // v_{I+1} is one product away
  __shared__ __attribute__((aligned(16))) T Hs[INV ? NB * kInvLd : 4];
  if (stop && *stop) return;  // every workgroup reads the same flag
  const WaitGuard wg(status, wait_ticks);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int LAST = INV ? 1 : 0;  // blocks applied after the main loop
  for (int I = nblk - 1 - blockIdx.x; I >= 0; I -= gridDim.x) {
    const int i0 = I * NB;
    const int in = min(NB, n - i0);
    const int rA = i0 + lane, rB = i0 + H + lane;
    const bool okA = lane < in, okB = H + lane < in;
    // permutation entries of the rows this thread publishes, read now so
    // that no global load sits between the block's solution and its records
    int pub_perm = 0;
    if constexpr (INV) {
      if (threadIdx.x < in) pub_perm = perm[i0 + threadIdx.x];
    } else {
      if (w == 0 && okA) pub_perm = perm[rA];
      if (w == 1 && okB) pub_perm = perm[rB];
    }
    if constexpr (INV) {
      stage_inverse_block<T, true>(Lp, Li + (int64_t)I * NB * NB);
      if (I + 1 < nblk) stage_inverse_block<T, false>(Hs, Li + (int64_t)(2 * nblk + I) * NB * NB);
    } else {
      stage_diag_block<T>(Lp, F, lda, piv, i0, in, n);
    }
    T wv[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) wv[q] = T(0);
    if (w == 0 && okA) dsolve_row<T, NR>(F, lda, piv, Z, ldz, rA, wv);
    if (w == 1 && okB) dsolve_row<T, NR>(F, lda, piv, Z, ldz, rB, wv);
    // 2x2 pivot across the block edge: L(i0 + 128, i0 + 127) = 0 (row
    // i0 + 127 is lane 63's row B; column i0 + 128 is the first column of
    // block I + 1: wavefront 0, first chunk, first column)
    const bool edge2 = w == 0 && lane == H - 1 && okB && I + 1 < nblk && piv[rB] == 2;
    T acc[2][NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[0][q] = acc[1][q] = T(0);
    constexpr int UB = sizeof(T) == 8 ? 1 : 2;
    const int64_t roA = (int64_t)min(rA, n - 1) * lda, roB = (int64_t)min(rB, n - 1) * lda;
    // Each lane reads 16 consecutive columns of its two rows: as 16-byte
    // vectors when the rows are 16-byte aligned (one address per 16 bytes
    // instead of per element: the rows differ per lane, so every load
    // instruction touches 64 cache lines)
    constexpr int VE = 16 / sizeof(T);
    typedef T vec_t __attribute__((ext_vector_type(VE)));
    const bool vec_ok = lda % VE == 0 && (reinterpret_cast<uintptr_t>(F) & 15) == 0;
    const int Im = I + LAST;  // blocks J > Im in the main loop
    for (int J = nblk - 1; J > Im; J -= UB) {
      T fa[2 * UB][16], fb[2 * UB][16];
      int row0[2 * UB];
#pragma unroll
      for (int k = 0; k < 2 * UB; ++k) {
        const int Ju = J - k / 2;
        const int cb = max(Ju, Im + 1) * NB + 32 * w + 16 * (k & 1);
        if constexpr (INV) {
          const int rAc = min(rA, n - 1), rBc = min(rB, n - 1);
#pragma unroll
          for (int cc = 0; cc < 16; ++cc) {
            const int64_t co = (int64_t)min(cb + cc, n - 1) * lda;
            fa[k][cc] = F[co + rAc];
            fb[k][cc] = F[co + rBc];
          }
        } else if (vec_ok && cb + 16 <= n) {
          const vec_t* pa = reinterpret_cast<const vec_t*>(F + roA + cb);
          const vec_t* pb = reinterpret_cast<const vec_t*>(F + roB + cb);
#pragma unroll
          for (int u = 0; u < 16 / VE; ++u) {
            const vec_t a = pa[u], b = pb[u];
#pragma unroll
            for (int e = 0; e < VE; ++e) {
              fa[k][u * VE + e] = a[e];
              fb[k][u * VE + e] = b[e];
            }
          }
        } else {
#pragma unroll
          for (int cc = 0; cc < 16; ++cc) {
            const int col = min(cb + cc, n - 1);
            fa[k][cc] = F[roA + col];
            fb[k][cc] = F[roB + col];
          }
        }
        row0[k] = max(Ju, Im + 1) * NB + 32 * w + 16 * (k & 1);
      }
#pragma unroll
      for (int k = 0; k < 2 * UB; ++k) {
        const int Ju = J - k / 2;
        const int cb = Ju * NB + 32 * w + 16 * (k & 1);
        const bool live = Ju > Im;
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) {
          const bool colok = live && cb + cc < n;
          if (!(colok && okA)) fa[k][cc] = T(0);
          if (!(colok && okB)) fb[k][cc] = T(0);
        }
        if (edge2 && Ju == I + 1 && (k & 1) == 0) fb[k][0] = T(0);
      }
      double vv[2 * UB];
      poll_rows_multi<NR, 2 * UB>(rec, row0, n, tag, vv, wg);
#pragma unroll
      for (int k = 0; k < 2 * UB; ++k)
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) {
#pragma unroll
          for (int q = 0; q < NR; ++q) {
            const T v = static_cast<T>(bcast(vv[k], cc * 4 + q));
            acc[0][q] += fa[k][cc] * v;
            acc[1][q] += fb[k][cc] * v;
          }
        }
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      part[w][lane][q] = acc[0][q];
      part[w][H + lane][q] = acc[1][q];
    }
    __syncthreads();
    if constexpr (INV) {
      // s' into vbot (rows 0..127), then u = Li^T s'
      if (w < 2) {
        const int r = w * H + lane;
        const bool ok = w == 0 ? okA : okB;
#pragma unroll
        for (int q = 0; q < NR; ++q)
          vbot[r][q] = ok ? wv[q] - ((part[0][r][q] + part[1][r][q]) +
                                     (part[2][r][q] + part[3][r][q]))
                          : T(0);
      }
      __syncthreads();
      T sv[NR];
      inverse_product<T, NR>(Lp, vbot, sv);
      const int r = threadIdx.x & (NB - 1);
      if (threadIdx.x >= NB) {
#pragma unroll
        for (int q = 0; q < NR; ++q) part[0][r][q] = sv[q];
      }
      __syncthreads();
      T u[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q) u[q] = sv[q] + part[0][r][q];
      if (I + 1 < nblk) {
        // block I+1's solution: wavefront w polls rows 32w .. 32w+31 (0
        // past n) into vbot (free again), then u -= H_I v_{I+1}
        const int cb = (I + 1) * NB + 32 * w;
        const int row0[2] = {cb, cb + 16};
        double vv[2];
        poll_rows_multi<NR, 2>(rec, row0, n, tag, vv, wg);
        if ((lane & 3) < NR) {
#pragma unroll
          for (int k = 0; k < 2; ++k)
            vbot[32 * w + 16 * k + (lane >> 2)][lane & 3] = static_cast<T>(vv[k]);
        }
        __syncthreads();
        T gv[NR];
        inverse_product<T, NR>(Hs, vbot, gv);
        if (threadIdx.x >= NB) {
#pragma unroll
          for (int q = 0; q < NR; ++q) part[1][r][q] = gv[q];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NR; ++q) u[q] -= gv[q] + part[1][r][q];
      }
      if (threadIdx.x < NB && r < in) {
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          xput(rec + (int64_t)(i0 + r) * kSolveMaxRhs + q, tag, q, static_cast<double>(u[q]));
          X[(int64_t)pub_perm * ldx + q] = u[q];
        }
      }
      __syncthreads();  // Lp, Hs, part, vbot are reused by the next block
      continue;
    }
    if (w == 1) {
      T v[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q)
        v[q] = okB ? wv[q] - ((part[0][H + lane][q] + part[1][H + lane][q]) +
                              (part[2][H + lane][q] + part[3][H + lane][q]))
                   : T(0);
      const int r = H + lane;  // L(c, r) at Lp[c (c - 1) / 2 + r], c = H + cc > r
      serial64<T, NR, false>(v, Lp, lane, [&](int cc) {
        return min((H + cc) * (H + cc - 1) / 2 + r, NPK - 1);
      });
#pragma unroll
      for (int q = 0; q < NR; ++q) vbot[lane][q] = okB ? v[q] : T(0);
      if (okB) {
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          xput(rec + (int64_t)rB * kSolveMaxRhs + q, tag, q, static_cast<double>(v[q]));
          X[(int64_t)pub_perm * ldx + q] = v[q];
        }
      }
    }
    __syncthreads();
    if (w == 0) {
      T v[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q)
        v[q] = okA ? wv[q] - ((part[0][lane][q] + part[1][lane][q]) +
                              (part[2][lane][q] + part[3][lane][q]))
                   : T(0);
      // the bottom unknowns applied in chunks of 8 (loads ahead of products)
#pragma unroll
      for (int c0 = 0; c0 < H; c0 += 8) {
        T lc[8], vc[8][NR];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int c = H + c0 + u;
          lc[u] = Lp[c * (c - 1) / 2 + lane];
#pragma unroll
          for (int q = 0; q < NR; ++q) vc[u][q] = vbot[c0 + u][q];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int q = 0; q < NR; ++q) v[q] -= lc[u] * vc[u][q];
      }
      // L(c, lane) at Lp[c (c - 1) / 2 + lane], c > lane (c = 0 never
      // contributes)
      serial64<T, NR, false>(v, Lp, lane, [&](int c) {
        return min(c * (c - 1) / 2 + lane, NPK - 1);
      });
      if (okA) {
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          xput(rec + (int64_t)rA * kSolveMaxRhs + q, tag, q, static_cast<double>(v[q]));
          X[(int64_t)pub_perm * ldx + q] = v[q];
        }
      }
    }
    __syncthreads();  // Lp, part, vbot are reused by the next block
  }
}

// ---- binary64 residual / backward error --------------------------------
// Tile (I, J), I <= J, of the upper triangle: row sums sum_j A(i,j) x_j
// into prow[J][i] and column sums sum_{i<j} A(i,j) x_i into pcol[I][j]
// (strictly upper part only), plus the same with |A| and x = 1 when
// with_norm.  Reduced in a fixed order by k_symv_reduce.
constexpr int kSymvTile = 64;

// One butterfly step of a wavefront reduce-scatter: lanes l and l ^ D
// exchange the halves of v[0 .. 2H) they do not keep; lane l keeps the upper
// half when bit D of l is set.  Sums are formed in a fixed order.
template <int H, int D>
__device__ __forceinline__ void reduce_scatter_step(double* v, int lane) {
  const bool hi = (lane & D) != 0;
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const double send = hi ? v[k] : v[k + H];
    const double keep = hi ? v[k + H] : v[k];
    v[k] = keep + __shfl_xor(send, D);
  }
}

// Sum over the 64 lanes of f(u) for the 16 values u = 0 .. 15 of every
// lane: lane l returns the sum for u = (l >> 2) & 15 (all four lanes of a
// quad agree).  The first step forms f on the fly, so only eight partial
// sums are live.
template <typename F>
__device__ __forceinline__ double wave_sum16_scatter(F f, int lane) {
  double v[8];
  {
    const bool hi = (lane & 32) != 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double a = f(k), b = f(k + 8);
      const double send = hi ? a : b;
      const double keep = hi ? b : a;
      v[k] = keep + __shfl_xor(send, 32);
    }
  }
  reduce_scatter_step<4, 16>(v, lane);
  reduce_scatter_step<2, 8>(v, lane);
  reduce_scatter_step<1, 4>(v, lane);
  double x = v[0];
  x += dpp_f64<kDppQuadXor1, 0xf>(x);
  x += dpp_f64<kDppQuadXor2, 0xf>(x);
  return x;
}

// Wavefront w holds rows w + 4 u (u = 0 .. 15) of the tile, lane c column c,
// in registers: column sums are per-lane sums combined over the four
// wavefronts through LDS; row sums are a reduce-scatter across the lanes.
// The tile never passes through LDS, so six workgroups fit on a CU.  NR:
// right-hand sides (1 .. kSolveMaxRhs), a template argument so that every
// load is straight-line code and all of them are in flight together.
template <int NR>
__global__ __launch_bounds__(256) void k_symv_tiles(
    int n, const double* __restrict__ A, int64_t lda,
    const double* __restrict__ X, int64_t ldx, double* __restrict__ prow,
    double* __restrict__ pcol, double* __restrict__ pabs_row,
    double* __restrict__ pabs_col, int with_norm, const int* __restrict__ stop) {
  constexpr int TB = kSymvTile;
  constexpr int PER = TB / 4;
  static_assert(PER == 16, "wave_sum16_scatter reduces 16 rows per wavefront");
  if (stop && *stop) return;
  __shared__ double cpart[4][NR + 1][TB];
  __shared__ double xs_i[NR][TB];
  int idx = blockIdx.x;
  int J = static_cast<int>((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
  while (J * (J + 1) / 2 > idx) --J;
  while ((J + 1) * (J + 2) / 2 <= idx) ++J;
  const int I = idx - J * (J + 1) / 2;
  const int i0 = I * TB, j0 = J * TB;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int jj = j0 + lane;
  const int jc = min(jj, n - 1);
  // every load issued before the first wait: clamped addresses, masked after
  double xj[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) xj[q] = X[(int64_t)jc * ldx + q];
  double xi = 0.0;
  const int xr = t / NR, xq = t % NR;
  if (t < TB * NR) xi = X[(int64_t)min(i0 + xr, n - 1) * ldx + xq];
  double v[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int i = min(i0 + w + 4 * u, n - 1);
    v[u] = A[(int64_t)i * lda + jc];
  }
  if (t < TB * NR) xs_i[xq][xr] = i0 + xr < n ? xi : 0.0;
#pragma unroll
  for (int q = 0; q < NR; ++q) xj[q] = jj < n ? xj[q] : 0.0;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int i = i0 + w + 4 * u;
    if (!(i < n && jj < n && jj >= i)) v[u] = 0.0;
  }
  __syncthreads();  // xs_i
  // column sums over the strictly upper part
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < PER; ++u) s += i0 + w + 4 * u < jj ? v[u] * xs_i[q][w + 4 * u] : 0.0;
    cpart[w][q][lane] = s;
  }
  if (with_norm) {
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < PER; ++u) s += i0 + w + 4 * u < jj ? fabs(v[u]) : 0.0;
    cpart[w][NR][lane] = s;
  }
  // row sums (diagonal included)
  const int irow = i0 + w + 4 * ((lane >> 2) & 15);
  const bool wr = (lane & 3) == 0 && irow < n;
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const double x = xj[q];
    const double s = wave_sum16_scatter([&](int u) { return v[u] * x; }, lane);
    if (wr) prow[((int64_t)J * n + irow) * NR + q] = s;
  }
  if (with_norm) {
    const double s = wave_sum16_scatter([&](int u) { return fabs(v[u]); }, lane);
    if (wr) pabs_row[(int64_t)J * n + irow] = s;
  }
  __syncthreads();
  if (w != 0 || jj >= n) return;
#pragma unroll
  for (int q = 0; q <= NR; ++q) {
    if (q == NR && !with_norm) break;
    const double c = ((cpart[0][q][lane] + cpart[1][q][lane]) + cpart[2][q][lane]) + cpart[3][q][lane];
    if (q < NR)
      pcol[((int64_t)I * n + jj) * NR + q] = c;
    else
      pabs_col[(int64_t)I * n + jj] = c;
  }
}

__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
  // non-negative binary64 values order like their bit patterns
  if (!(v >= 0.0)) v = __builtin_huge_val();  // NaN poisons the norm
  atomicMax(reinterpret_cast<unsigned long long*>(addr),
            __double_as_longlong(v));
}

// y_i = sum_{J >= I(i)} prow[J][i] + sum_{I' <= J(i)} pcol[I'][i] in a fixed
// order: 64 rows per workgroup, four wavefronts each summing every fourth
// tile's partial, combined in wavefront order; R = B - Y (binary64); norms
// into norms[0..]:
//   norms[q] = max|r_q|, norms[4+q] = max|x_q|, norms[8+q] = max|b_q|,
//   norms[12] = ||A||_inf (when with_norm).
__global__ __launch_bounds__(256) void k_symv_reduce(
    int n, int nr, const double* __restrict__ prow,
    const double* __restrict__ pcol, const double* __restrict__ pabs_row,
    const double* __restrict__ pabs_col, const double* __restrict__ Bv,
    int64_t ldb, const double* __restrict__ X, int64_t ldx,
    double* __restrict__ R, int64_t ldr, double* __restrict__ norms,
    int with_norm, const int* __restrict__ stop) {
  __shared__ double part[4][64][kSolveMaxRhs + 1];
  if (stop && *stop) return;
  const int li = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + li;
  const bool ok = i < n;
  const int T = (n + kSymvTile - 1) / kSymvTile;
  const int ti = ok ? i / kSymvTile : 0;
  double acc[kSolveMaxRhs + 1] = {0, 0, 0, 0, 0};
  if (ok) {
#pragma unroll 4
    for (int J = ti + w; J < T; J += 4)
      for (int q = 0; q < nr; ++q) acc[q] += prow[((int64_t)J * n + i) * nr + q];
#pragma unroll 4
    for (int I = w; I <= ti; I += 4)
      for (int q = 0; q < nr; ++q) acc[q] += pcol[((int64_t)I * n + i) * nr + q];
    if (with_norm) {
#pragma unroll 4
      for (int J = ti + w; J < T; J += 4) acc[kSolveMaxRhs] += pabs_row[(int64_t)J * n + i];
#pragma unroll 4
      for (int I = w; I <= ti; I += 4) acc[kSolveMaxRhs] += pabs_col[(int64_t)I * n + i];
    }
  }
  for (int q = 0; q <= kSolveMaxRhs; ++q) part[w][li][q] = acc[q];
  __syncthreads();
  if (w != 0 || !ok) return;
  for (int q = 0; q < nr; ++q) {
    const double y = ((part[0][li][q] + part[1][li][q]) + part[2][li][q]) + part[3][li][q];
    const double b = Bv[(int64_t)i * ldb + q];
    const double r = b - y;
    if (R) R[(int64_t)i * ldr + q] = r;
    atomic_max_nonneg(&norms[q], fabs(r));
    atomic_max_nonneg(&norms[4 + q], fabs(X[(int64_t)i * ldx + q]));
    atomic_max_nonneg(&norms[8 + q], fabs(b));
  }
  if (with_norm) {
    const double sabs = ((part[0][li][kSolveMaxRhs] + part[1][li][kSolveMaxRhs]) +
                         part[2][li][kSolveMaxRhs]) + part[3][li][kSolveMaxRhs];
    atomic_max_nonneg(&norms[12], sabs);
  }
}

__global__ void k_berr(int nr, const double* __restrict__ norms,
                       double anorm_override, double* __restrict__ eta) {
  const int q = threadIdx.x;
  if (q >= nr) return;
  const double anorm = anorm_override >= 0.0 ? anorm_override : norms[12];
  const double den = anorm * norms[4 + q] + norms[8 + q];
  eta[q] = den > 0.0 ? norms[q] / den : (norms[q] > 0.0 ? __builtin_huge_val() : 0.0);
}

// Refinement stopping test on the device, after correction number it:
// *best is the smallest max_q eta_q so far, ctl[2] = 1 when this correction
// improved it (else 0); stop (ctl[0] = 1, ctl[1] = it) when the largest
// backward error is <= tol, is NaN, or it reached max_iter.  Once set, the
// flag stays: later checks and the kernels that take it do nothing.
__global__ void k_refine_check(const double* __restrict__ eta, int nr, double tol, int it,
                               int max_iter, int* __restrict__ ctl, double* __restrict__ best) {
  if (threadIdx.x != 0 || ctl[0]) return;
  double m = 0.0;
  bool nan = false;
  for (int q = 0; q < nr; ++q) {
    const double e = eta[q];
    if (!(e == e)) nan = true;
    m = e > m ? e : m;
  }
  const bool improved = !nan && m < *best;
  if (improved) *best = m;
  ctl[2] = improved ? 1 : 0;
  if (nan || m <= tol || it >= max_iter) {
    ctl[0] = 1;
    ctl[1] = it;
  }
}

// Xb := X (n x nr, Xb with leading dimension nr) when the last stopping
// test recorded an improvement, so that Xb holds the best iterate.
__global__ void k_keep_best(int n, int nr, const double* __restrict__ X, int64_t ldx,
                            double* __restrict__ Xb, const int* __restrict__ ctl) {
  if (!ctl[2]) return;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * nr) return;
  const int64_t i = e / nr, q = e % nr;
  Xb[i * nr + q] = X[i * ldx + q];
}

__global__ void k_fill_f64(double* p, int count, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}
__global__ void k_fill_i32(int* p, int count, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

// Upper triangle (j >= i) of a row-major n x n matrix, converted S -> D
// (a copy when S = D): one workgroup per row, four consecutive elements per
// thread and step, as 16- or 32-byte vectors when VEC (leading dimensions
// multiples of 4, rows aligned to four elements).  The lower triangle of
// dst is not written.
template <typename S, typename D, bool VEC>
__global__ __launch_bounds__(256) void k_copy_upper(int n, const S* __restrict__ src, int64_t lds,
                                                    D* __restrict__ dst, int64_t ldd) {
  const int i = blockIdx.x;
  const S* sr = src + (int64_t)i * lds;
  D* dr = dst + (int64_t)i * ldd;
  if constexpr (VEC) {
    typedef S s4 __attribute__((ext_vector_type(4)));
    typedef D d4 __attribute__((ext_vector_type(4)));
    for (int j = (i & ~3) + 4 * (int)threadIdx.x; j < n; j += 4 * 256) {
      if (j >= i && j + 4 <= n) {
        const s4 v = *reinterpret_cast<const s4*>(sr + j);
        d4 o;
        o.x = static_cast<D>(v.x);
        o.y = static_cast<D>(v.y);
        o.z = static_cast<D>(v.z);
        o.w = static_cast<D>(v.w);
        *reinterpret_cast<d4*>(dr + j) = o;
      } else {
        for (int t = 0; t < 4; ++t)
          if (j + t >= i && j + t < n) dr[j + t] = static_cast<D>(sr[j + t]);
      }
    }
  } else {
    for (int j = i + (int)threadIdx.x; j < n; j += 256) dr[j] = static_cast<D>(sr[j]);
  }
}

template <typename S, typename D>
static void copy_upper(Handle* h, int n, const S* src, int64_t lds, D* dst, int64_t ldd) {
  const bool vec = lds % 4 == 0 && ldd % 4 == 0 &&
                   reinterpret_cast<uintptr_t>(src) % (4 * sizeof(S)) == 0 &&
                   reinterpret_cast<uintptr_t>(dst) % (4 * sizeof(D)) == 0;
  if (vec)
    hipLaunchKernelGGL((k_copy_upper<S, D, true>), dim3(n), dim3(256), 0, h->stream, n, src, lds,
                       dst, ldd);
  else
    hipLaunchKernelGGL((k_copy_upper<S, D, false>), dim3(n), dim3(256), 0, h->stream, n, src, lds,
                       dst, ldd);
}

template <typename S, typename D>
__global__ void k_copy_cols(int n, int nr, const S* __restrict__ src,
                            int64_t lds, D* __restrict__ dst, int64_t ldd) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * nr) return;
  const int i = e / nr, q = e % nr;
  dst[(int64_t)i * ldd + q] = (D)src[(int64_t)i * lds + q];
}

// X += (double) Dv, unless *stop
__global__ void k_axpy_f32_to_f64(int n, int nr, const float* __restrict__ dv,
                                  int64_t ldd, double* __restrict__ X,
                                  int64_t ldx, const int* __restrict__ stop) {
  if (*stop) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * nr) return;
  const int i = e / nr, q = e % nr;
  X[(int64_t)i * ldx + q] += (double)dv[(int64_t)i * ldd + q];
}

static unsigned blocks_for(int64_t count, int threads) {
  return static_cast<unsigned>((count + threads - 1) / threads);
}

template <typename K>
static int coop_groups(K kernel, int want) {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kernel),
                                                   kSolveThreads, 0) != hipSuccess)
    return 1;
  if (hipGetDevice(&dev) != hipSuccess) return 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 1;
  return std::max(1, std::min(want, per_cu * cus));
}

// Scratch for one solve: records for both sweeps (n x kSolveMaxRhs each)
// and Z (n x kSolveMaxRhs of T).  The record area is zeroed when it is
// (re)allocated; afterwards every launch uses a fresh tag, so records of
// earlier launches are never accepted and nothing is cleared per call.
template <typename T, int NR, bool INV>
static int solve_launch(Handle* h, int n, int nblk, const T* F, int64_t lda, const int* perm,
                        const int* piv, const T* B, int64_t ldb, T* X, int64_t ldx, T* Z,
                        XRec* rf, XRec* rb, const T* Li, const int* stop) {
  const int gf = coop_groups(k_trsv_fwd<T, NR, INV>, nblk);
  const int gb = coop_groups(k_trsv_bwd<T, NR, INV>, nblk);
  int64_t ldz = kSolveMaxRhs;
  {
    int tag = h->next_solve_tag();
    int n_ = n, nblk_ = nblk;
    int64_t lda_ = lda, ldb_ = ldb;
    const T* F_ = F;
    const int *piv_ = piv, *perm_ = perm;
    const T* B_ = B;
    T* Z_ = Z;
    XRec* r_ = rf;
    const T* Li_ = Li;
    const int* stop_ = stop;
    unsigned* st_ = h->d_status;
    long long pl_ = h->wait_ticks();
    void* args[] = {&n_,  &F_,   &lda_, &piv_,  &perm_, &B_,   &ldb_, &Z_,
                    &ldz, &r_,   &tag,  &nblk_, &Li_,   &stop_, &st_, &pl_};
    hipError_t e = launch_resident(reinterpret_cast<const void*>(&k_trsv_fwd<T, NR, INV>),
                                   dim3(gf), dim3(kSolveThreads), args, h->stream, true);
    if (e != hipSuccess) return h->hip_fail(e, "forward solve launch");
  }
  {
    int tag = h->next_solve_tag();
    int n_ = n, nblk_ = nblk;
    int64_t lda_ = lda, ldx_ = ldx;
    const T* F_ = F;
    const int *piv_ = piv, *perm_ = perm;
    const T* Z_ = Z;
    T* X_ = X;
    XRec* r_ = rb;
    const T* Li_ = Li;
    const int* stop_ = stop;
    unsigned* st_ = h->d_status;
    long long pl_ = h->wait_ticks();
    void* args[] = {&n_,  &F_,   &lda_, &piv_,  &perm_, &Z_,   &ldz,  &X_,
                    &ldx_, &r_,  &tag,  &nblk_, &Li_,   &stop_, &st_, &pl_};
    hipError_t e = launch_resident(reinterpret_cast<const void*>(&k_trsv_bwd<T, NR, INV>),
                                   dim3(gb), dim3(kSolveThreads), args, h->stream, true);
    if (e != hipSuccess) return h->hip_fail(e, "backward solve launch");
  }
  return BBK_OK;
}

// Li: the diagonal blocks' inverses from k_diag_inv followed by the
// forward and backward sweeps' next-block products (k_next_block): 3 nblk
// blocks of 128 x 128, binary32 only; or NULL for substitution.
template <typename T>
int solve(Handle* h, int n, int nrhs, const T* F, int64_t lda, const int* perm,
          const int* piv, T* B, int64_t ldb, const T* Li, const int* stop) {
  if (n <= 0) return -2;
  if (nrhs < 0) return -3;
  if (!F) return -4;
  if (lda < n) return -5;
  if (!perm) return -6;
  if (!piv) return -7;
  if (!B) return -8;
  if (ldb < std::max(1, nrhs)) return -9;
  if (nrhs == 0) return BBK_OK;
  const int nblk = (n + kSolveBlock - 1) / kSolveBlock;
  const size_t rec_bytes = (size_t)n * kSolveMaxRhs * sizeof(XRec);
  const size_t vec_bytes = ((size_t)n * kSolveMaxRhs * sizeof(T) + 255) / 256 * 256;
  int rc = h->reserve_solve(2 * rec_bytes + vec_bytes);
  if (rc != BBK_OK) return rc;
  char* base = reinterpret_cast<char*>(h->solve_ws);
  XRec* rf = reinterpret_cast<XRec*>(base);
  XRec* rb = reinterpret_cast<XRec*>(base + rec_bytes);
  T* Z = reinterpret_cast<T*>(base + 2 * rec_bytes);
  // Record areas are cleared when the buffer was (re)allocated (the same
  // address may come back with other contents), when the order changes
  // the layout (rb and Z move with n), or when the tags would wrap.
  if (h->solve_gen != h->solve_clear_gen || h->solve_layout != n || h->solve_tag > (1 << 30)) {
    const int words = static_cast<int>(2 * rec_bytes / sizeof(int));
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(words, 256)), dim3(256), 0, h->stream,
                       reinterpret_cast<int*>(base), words, 0);
    h->solve_clear_gen = h->solve_gen;
    h->solve_layout = n;
    h->solve_tag = 0;
  }
  for (int q0 = 0; q0 < nrhs; q0 += kSolveMaxRhs) {
    const int nr = std::min(kSolveMaxRhs, nrhs - q0);
    const T* Bq = B + q0;
    T* Xq = B + q0;
    if constexpr (std::is_same<T, float>::value) {
      if (Li) {
        switch (nr) {
          case 1: rc = solve_launch<T, 1, true>(h, n, nblk, F, lda, perm, piv, Bq, ldb, Xq, ldb, Z, rf, rb, Li, stop); break;
          case 2: rc = solve_launch<T, 2, true>(h, n, nblk, F, lda, perm, piv, Bq, ldb, Xq, ldb, Z, rf, rb, Li, stop); break;
          case 3: rc = solve_launch<T, 3, true>(h, n, nblk, F, lda, perm, piv, Bq, ldb, Xq, ldb, Z, rf, rb, Li, stop); break;
          default: rc = solve_launch<T, 4, true>(h, n, nblk, F, lda, perm, piv, Bq, ldb, Xq, ldb, Z, rf, rb, Li, stop); break;
        }
        if (rc != BBK_OK) return rc;
        continue;
      }
    }
    switch (nr) {
      case 1: rc = solve_launch<T, 1, false>(h, n, nblk, F, lda, perm, piv, Bq, ldb, Xq, ldb, Z, rf, rb, nullptr, stop); break;
      case 2: rc = solve_launch<T, 2, false>(h, n, nblk, F, lda, perm, piv, Bq, ldb, Xq, ldb, Z, rf, rb, nullptr, stop); break;
      case 3: rc = solve_launch<T, 3, false>(h, n, nblk, F, lda, perm, piv, Bq, ldb, Xq, ldb, Z, rf, rb, nullptr, stop); break;
      default: rc = solve_launch<T, 4, false>(h, n, nblk, F, lda, perm, piv, Bq, ldb, Xq, ldb, Z, rf, rb, nullptr, stop); break;
    }
    if (rc != BBK_OK) return rc;
  }
  return BBK_OK;
}

// Backward error of X for A X = B (binary64).  R (may be NULL) receives
// B - A X.  anorm: pass < 0 to compute ||A||_inf (written to *anorm_out).
int berr(Handle* h, int n, int nrhs, const double* A, int64_t lda,
         const double* B, int64_t ldb, const double* X, int64_t ldx, double* R,
         int64_t ldr, double* d_eta, double anorm, bool want_norm, const int* stop) {
  const int T = (n + kSymvTile - 1) / kSymvTile;
  const size_t part = (size_t)T * n * kSolveMaxRhs * sizeof(double);
  const size_t partabs = (size_t)T * n * sizeof(double);
  const size_t nbytes = 256;
  int rc = h->reserve_berr(2 * part + 2 * partabs + nbytes);
  if (rc != BBK_OK) return rc;
  char* base = reinterpret_cast<char*>(h->berr_ws);
  double* prow = reinterpret_cast<double*>(base);
  double* pcol = reinterpret_cast<double*>(base + part);
  double* parow = reinterpret_cast<double*>(base + 2 * part);
  double* pacol = reinterpret_cast<double*>(base + 2 * part + partabs);
  double* norms = reinterpret_cast<double*>(base + 2 * part + 2 * partabs);
  const int with_norm = want_norm ? 1 : 0;
  for (int q0 = 0; q0 < nrhs; q0 += kSolveMaxRhs) {
    const int nr = std::min(kSolveMaxRhs, nrhs - q0);
    hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(32), 0, h->stream, norms, 16, 0.0);
    const int tiles = T * (T + 1) / 2;
    switch (nr) {
      case 1: hipLaunchKernelGGL(k_symv_tiles<1>, dim3(tiles), dim3(256), 0, h->stream, n, A, lda, X + q0, ldx, prow, pcol, parow, pacol, with_norm, stop); break;
      case 2: hipLaunchKernelGGL(k_symv_tiles<2>, dim3(tiles), dim3(256), 0, h->stream, n, A, lda, X + q0, ldx, prow, pcol, parow, pacol, with_norm, stop); break;
      case 3: hipLaunchKernelGGL(k_symv_tiles<3>, dim3(tiles), dim3(256), 0, h->stream, n, A, lda, X + q0, ldx, prow, pcol, parow, pacol, with_norm, stop); break;
      default: hipLaunchKernelGGL(k_symv_tiles<4>, dim3(tiles), dim3(256), 0, h->stream, n, A, lda, X + q0, ldx, prow, pcol, parow, pacol, with_norm, stop); break;
    }
    hipLaunchKernelGGL(k_symv_reduce, dim3(blocks_for(n, 64)), dim3(256), 0, h->stream, n, nr,
                       prow, pcol, parow, pacol, B + q0, ldb, X + q0, ldx,
                       R ? R + q0 : nullptr, ldr, norms, with_norm, stop);
    hipLaunchKernelGGL(k_berr, dim3(1), dim3(32), 0, h->stream, nr, norms,
                       want_norm ? -1.0 : anorm, d_eta + q0);
    if (want_norm && q0 == 0) {
      // keep ||A||_inf for later calls (copied by the caller if needed)
      h->last_anorm_dev = norms + 12;
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return h->hip_fail(e, "backward error kernels");
  return BBK_OK;
}

template int solve<double>(Handle*, int, int, const double*, int64_t, const int*, const int*,
                           double*, int64_t, const double*, const int*);
template int solve<float>(Handle*, int, int, const float*, int64_t, const int*, const int*,
                          float*, int64_t, const float*, const int*);

}  // namespace bbk

using bbk::Handle;

extern "C" {

int bbk_dsytrs(bbk_handle_t handle, int n, int nrhs, const double* d_F, int64_t lda,
               const int* d_perm, const int* d_piv, double* d_B, int64_t ldb) {
  if (!handle) return -1;
  return bbk::solve<double>(reinterpret_cast<Handle*>(handle), n, nrhs, d_F, lda, d_perm,
                            d_piv, d_B, ldb);
}

int bbk_ssytrs(bbk_handle_t handle, int n, int nrhs, const float* d_F, int64_t lda,
               const int* d_perm, const int* d_piv, float* d_B, int64_t ldb) {
  if (!handle) return -1;
  return bbk::solve<float>(reinterpret_cast<Handle*>(handle), n, nrhs, d_F, lda, d_perm,
                           d_piv, d_B, ldb);
}

int bbk_dsyberr(bbk_handle_t handle, int n, int nrhs, const double* d_A, int64_t lda,
                const double* d_B, int64_t ldb, const double* d_X, int64_t ldx,
                double* d_eta) {
  if (!handle) return -1;
  if (n <= 0) return -2;
  if (nrhs < 0) return -3;
  if (!d_A) return -4;
  if (lda < n) return -5;
  if (!d_B) return -6;
  if (ldb < std::max(1, nrhs)) return -7;
  if (!d_X) return -8;
  if (ldx < std::max(1, nrhs)) return -9;
  if (!d_eta) return -10;
  if (nrhs == 0) return BBK_OK;
  Handle* h = reinterpret_cast<Handle*>(handle);
  return bbk::berr(h, n, nrhs, d_A, lda, d_B, ldb, d_X, ldx, nullptr, 0, d_eta, -1.0, true);
}

static int check_solve_args(int n, int nrhs, const void* A, int64_t lda, const void* B,
                            int64_t ldb, const void* X, int64_t ldx) {
  if (n <= 0) return -2;
  if (nrhs <= 0) return -3;
  if (!A) return -4;
  if (lda < n) return -5;
  if (!B) return -6;
  if (ldb < nrhs) return -7;
  if (!X) return -8;
  if (ldx < nrhs) return -9;
  return BBK_OK;
}

int bbk_dsysv(bbk_handle_t handle, int n, int nrhs, const double* d_A, int64_t lda,
              const double* d_B, int64_t ldb, double* d_X, int64_t ldx, double* berr_out) {
  if (!handle) return -1;
  int rc = check_solve_args(n, nrhs, d_A, lda, d_B, ldb, d_X, ldx);
  if (rc != BBK_OK) return rc;
  Handle* h = reinterpret_cast<Handle*>(handle);
  int64_t ldf = (n + 63) / 64 * 64;
  if (ldf % 1024 == 0) ldf += 64;
  rc = h->reserve_mixed((size_t)ldf * n * sizeof(double) + 2 * (size_t)n * sizeof(int) +
                        (size_t)nrhs * sizeof(double) + 1024);
  if (rc != BBK_OK) return rc;
  char* base = reinterpret_cast<char*>(h->mixed_ws);
  double* F = reinterpret_cast<double*>(base);
  int* perm = reinterpret_cast<int*>(base + (size_t)ldf * n * sizeof(double));
  int* piv = perm + n;
  double* eta = reinterpret_cast<double*>(base + (size_t)ldf * n * sizeof(double) +
                                          (2 * (size_t)n * sizeof(int) + 255) / 256 * 256);
  // the factorization reads and writes the upper triangle only
  bbk::copy_upper<double, double>(h, n, d_A, lda, F, ldf);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return h->hip_fail(e, "copy A");
  rc = bbk::factorize<double>(h, n, 0, F, ldf, perm, piv);
  if (rc != BBK_OK) return rc;
  hipLaunchKernelGGL((bbk::k_copy_cols<double, double>), dim3(bbk::blocks_for((int64_t)n * nrhs, 256)),
                     dim3(256), 0, h->stream, n, nrhs, d_B, ldb, d_X, ldx);
  rc = bbk::solve<double>(h, n, nrhs, F, ldf, perm, piv, d_X, ldx);
  if (rc != BBK_OK) return rc;
  rc = bbk::berr(h, n, nrhs, d_A, lda, d_B, ldb, d_X, ldx, nullptr, 0, eta, -1.0, true);
  if (rc != BBK_OK) return rc;
  std::vector<double> he(nrhs);
  e = hipMemcpyAsync(he.data(), eta, nrhs * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return h->hip_fail(e, "read backward error");
  if (berr_out) *berr_out = *std::max_element(he.begin(), he.end());
  return h->check_status();
}

int bbk_dsysv_mixed(bbk_handle_t handle, int n, int nrhs, const double* d_A, int64_t lda,
                    const double* d_B, int64_t ldb, double* d_X, int64_t ldx, double tol,
                    int max_iter, int* iters, double* berr_out) {
  if (!handle) return -1;
  int rc = check_solve_args(n, nrhs, d_A, lda, d_B, ldb, d_X, ldx);
  if (rc != BBK_OK) return rc;
  if (max_iter < 0) return -11;
  Handle* h = reinterpret_cast<Handle*>(handle);
  int64_t ldf = (n + 63) / 64 * 64;
  if (ldf % 1024 == 0) ldf += 64;
  const size_t fbytes = ((size_t)ldf * n * sizeof(float) + 255) / 256 * 256;
  const size_t ibytes = (2 * (size_t)n * sizeof(int) + 255) / 256 * 256;
  const size_t rbytes = ((size_t)n * nrhs * sizeof(double) + 255) / 256 * 256;
  const size_t dbytes = ((size_t)n * nrhs * sizeof(float) + 255) / 256 * 256;
  const int nblk = (n + bbk::kSolveBlock - 1) / bbk::kSolveBlock;
  // the diagonal blocks' inverses, then the two sweeps' next-block products
  const size_t lbytes = 3 * (size_t)nblk * bbk::kSolveBlock * bbk::kSolveBlock * sizeof(float);
  // eta (nrhs), the best backward error and the stop flags
  const size_t ebytes = ((size_t)nrhs * sizeof(double) + 64 + 255) / 256 * 256;
  // the best iterate so far (n x nrhs)
  const size_t xbytes = rbytes;
  rc = h->reserve_mixed(fbytes + ibytes + rbytes + dbytes + ebytes + lbytes + xbytes);
  if (rc != BBK_OK) return rc;
  char* base = reinterpret_cast<char*>(h->mixed_ws);
  float* S = reinterpret_cast<float*>(base);
  int* perm = reinterpret_cast<int*>(base + fbytes);
  int* piv = perm + n;
  double* R = reinterpret_cast<double*>(base + fbytes + ibytes);
  float* Dv = reinterpret_cast<float*>(base + fbytes + ibytes + rbytes);
  double* eta = reinterpret_cast<double*>(base + fbytes + ibytes + rbytes + dbytes);
  double* best_dev = eta + nrhs;
  int* ctl = reinterpret_cast<int*>(best_dev + 1);
  float* Li = reinterpret_cast<float*>(base + fbytes + ibytes + rbytes + dbytes + ebytes);
  double* Xb = reinterpret_cast<double*>(base + fbytes + ibytes + rbytes + dbytes + ebytes + lbytes);
  // FP32 copy of the upper triangle, FP32 factorization
  bbk::copy_upper<double, float>(h, n, d_A, lda, S, ldf);
  rc = bbk::factorize<float>(h, n, 0, S, ldf, perm, piv);
  if (rc != BBK_OK) return rc;
  // inverses of the diagonal blocks of L: the refinement's solves multiply
  // by them instead of substituting; the backward sweeps read L^T from the
  // mirrored (otherwise unused) lower triangle of the binary32 factor
  hipLaunchKernelGGL(bbk::k_diag_inv<float>, dim3(nblk), dim3(bbk::kSolveBlock), 0, h->stream, n,
                     S, ldf, piv, Li);
  if (nblk > 1)
    hipLaunchKernelGGL(bbk::k_next_block<float>, dim3(nblk - 1, 2), dim3(4 * bbk::kSolveBlock), 0,
                       h->stream, n, S, ldf, piv, Li, nblk);
  {
    const unsigned tb = static_cast<unsigned>((n + 63) / 64);
    hipLaunchKernelGGL(bbk::k_mirror_upper<float>, dim3(tb * (tb + 1) / 2), dim3(256), 0,
                       h->stream, n, S, ldf);
  }
  // x0 = solve32(b)
  const unsigned eb = bbk::blocks_for((int64_t)n * nrhs, 256);
  hipLaunchKernelGGL((bbk::k_copy_cols<double, float>), dim3(eb), dim3(256), 0, h->stream, n,
                     nrhs, d_B, ldb, Dv, (int64_t)nrhs);
  rc = bbk::solve<float>(h, n, nrhs, S, ldf, perm, piv, Dv, nrhs, Li);
  if (rc != BBK_OK) return rc;
  hipLaunchKernelGGL((bbk::k_copy_cols<float, double>), dim3(eb), dim3(256), 0, h->stream, n,
                     nrhs, Dv, (int64_t)nrhs, d_X, ldx);
  // r = b - A x and eta in binary64 for x0, checked on the host (this call
  // also forms ||A||_inf)
  std::vector<double> he(nrhs);
  double anorm = -1.0;
  int it = 0;
  rc = bbk::berr(h, n, nrhs, d_A, lda, d_B, ldb, d_X, ldx, R, nrhs, eta, -1.0, true);
  if (rc != BBK_OK) return rc;
  hipError_t e = hipMemcpyAsync(&anorm, h->last_anorm_dev, sizeof(double),
                                hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(he.data(), eta, nrhs * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return h->hip_fail(e, "read backward error");
  double best = *std::max_element(he.begin(), he.end());
  for (double v : he)
    if (!(v == v)) best = v;
  if (!(best <= tol || it >= max_iter || !(best == best))) {
    // Corrections d = solve32(r), x += d, then r and eta again, queued a few
    // at a time with the stopping test on the device (k_refine_check): after
    // the test has stopped, the queued kernels return at once, so x is the
    // iterate the test accepted and the host waits once per chunk.  The
    // best iterate is kept (k_keep_best) and returned when the test stops
    // short of tol (max_iter reached, or a diverging correction).
    constexpr int kChunk = 4;
    hipLaunchKernelGGL(bbk::k_fill_i32, dim3(1), dim3(32), 0, h->stream, ctl, 3, 0);
    // x0 is the best iterate so far
    hipLaunchKernelGGL(bbk::k_fill_f64, dim3(1), dim3(32), 0, h->stream, best_dev, 1, best);
    hipLaunchKernelGGL((bbk::k_copy_cols<double, double>), dim3(eb), dim3(256), 0, h->stream, n,
                       nrhs, d_X, ldx, Xb, (int64_t)nrhs);
    int queued = 0;
    for (;;) {
      for (int c = 0; c < kChunk && queued < max_iter; ++c) {
        hipLaunchKernelGGL((bbk::k_copy_cols<double, float>), dim3(eb), dim3(256), 0, h->stream,
                           n, nrhs, R, (int64_t)nrhs, Dv, (int64_t)nrhs);
        rc = bbk::solve<float>(h, n, nrhs, S, ldf, perm, piv, Dv, nrhs, Li, ctl);
        if (rc != BBK_OK) return rc;
        hipLaunchKernelGGL(bbk::k_axpy_f32_to_f64, dim3(eb), dim3(256), 0, h->stream, n, nrhs,
                           Dv, (int64_t)nrhs, d_X, ldx, ctl);
        ++queued;
        rc = bbk::berr(h, n, nrhs, d_A, lda, d_B, ldb, d_X, ldx, R, nrhs, eta, anorm, false, ctl);
        if (rc != BBK_OK) return rc;
        hipLaunchKernelGGL(bbk::k_refine_check, dim3(1), dim3(32), 0, h->stream, eta, nrhs, tol,
                           queued, max_iter, ctl, best_dev);
        hipLaunchKernelGGL(bbk::k_keep_best, dim3(eb), dim3(256), 0, h->stream, n, nrhs, d_X, ldx,
                           Xb, ctl);
      }
      int hc[2] = {0, 0};
      e = hipMemcpyAsync(hc, ctl, sizeof(hc), hipMemcpyDeviceToHost, h->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(&best, best_dev, sizeof(double), hipMemcpyDeviceToHost, h->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
      if (e != hipSuccess) return h->hip_fail(e, "read refinement state");
      if (hc[0]) {
        it = hc[1];
        break;
      }
    }
    if (!(best <= tol)) {
      // not converged: return the best iterate, whose backward error is best
      hipLaunchKernelGGL((bbk::k_copy_cols<double, double>), dim3(eb), dim3(256), 0, h->stream,
                         n, nrhs, Xb, (int64_t)nrhs, d_X, ldx);
      e = hipStreamSynchronize(h->stream);
      if (e != hipSuccess) return h->hip_fail(e, "restore best iterate");
    }
  }
  if (iters) *iters = it;
  if (berr_out) *berr_out = best;
  rc = h->check_status();
  if (rc != BBK_OK) return rc;
  return best <= tol ? BBK_OK : BBK_ERR_NOT_CONVERGED;
}


}  // extern "C"

// bbk_update.hpp -- the MFMA trailing-update tile, shared by the
// single-matrix update kernel and the batched factorization.
//
// Trailing update C(i,j) -= sum_q U(s+q, i) W(q, j) over the upper
// triangle of [e, n)^2, e = s + width (UpdateTrailingBlocks +
// kernels::UpdateTrailingBlock, src/block_ldlt_decomposition.cc:706-744,
// src/kernels.cc:67-91).  Square output tiles per workgroup of 256
// threads, four wavefronts, 16x16x4 MFMA, K staged through LDS in chunks.
#pragma once

#include "bbk_internal.hpp"

namespace bbk {

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output element v of a lane sits in row (lane / 16) + 4 v (measured)
  static __device__ __forceinline__ int row(int lane, int v) {
    return (lane >> 4) + 4 * v;
  }
  static __device__ __forceinline__ int lane_row(int lane) { return lane >> 4; }
  static constexpr int v_row(int v) { return 4 * v; }
};
template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // output element v of a lane sits in row 4 (lane / 16) + v (measured)
  static __device__ __forceinline__ int row(int lane, int v) {
    return 4 * (lane >> 4) + v;
  }
  static __device__ __forceinline__ int lane_row(int lane) { return 4 * (lane >> 4); }
  static constexpr int v_row(int v) { return v; }
};

__device__ __forceinline__ void tri_tile(int idx, int* I, int* J) {
  int jj = static_cast<int>((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
  while (jj * (jj + 1) / 2 > idx) --jj;
  while ((jj + 1) * (jj + 2) / 2 <= idx) ++jj;
  *J = jj;
  *I = idx - jj * (jj + 1) / 2;
}

// The accumulators start from the C tile itself and the U operand is
// negated while it is staged, so the MFMA chain computes C - U^T W in
// place; K is staged through two LDS buffers so the next chunk's global
// loads overlap the current chunk's MFMAs.  TM x TM output tile per
// workgroup, four wavefronts of (TM/2) x (TM/2); MINB workgroups per CU;
// with XCD, consecutive tiles (sharing U and W strips) go to one XCD.
// Tile idx -> (I, J), I <= J < T, in 8 x 8 supertiles: supertile columns
// in order, within a column the full supertiles top to bottom, then the
// diagonal one; column-major inside a supertile.  A run of consecutive
// indices (one XCD's share) then reuses a few U and W strip pieces from L2
// instead of streaming whole strips.
__device__ __forceinline__ void supertile_of(int idx, int T, int* I, int* J) {
  constexpr int S = 8;
  int bj = 0, base = 0, C = 0;
  for (;;) {
    C = min(S, T - bj * S);
    const int col = bj * S * C + C * (C + 1) / 2;
    if (idx < base + col) break;
    base += col;
    ++bj;
  }
  const int l = idx - base;
  const int full = bj * S * C;
  if (l < full) {
    const int bi = l / (S * C), ll = l % (S * C);
    *I = bi * S + ll % S;
    *J = bj * S + ll / S;
  } else {
    int li, lj;
    tri_tile(l - full, &li, &lj);
    *I = bj * S + li;
    *J = bj * S + lj;
  }
}

// C tile element (ti, tj, v) of a lane: row wr WT + 16 ti + Mfma::row(lane, v),
// column wc WT + 16 tj + lane % 16 of the tile.  Its address is a
// wave-uniform part (tile origin, wavefront quadrant, ti, tj, v) plus one
// 32-bit lane offset, so the sixteen addresses live in scalar registers and
// the accumulators keep the vector registers.  FULL: a tile wholly above the
// diagonal and inside the matrix, loaded and stored without checks.
template <typename T, int TM, bool FULL>
__device__ __forceinline__ void c_tile_io(T* __restrict__ A, int64_t lda, int n, int i0, int j0,
                                          int wr, int wc, int lane,
                                          typename Mfma<T>::acc_t (*acc)[TM / 32], bool store) {
  constexpr int WT = TM / 2, TI = WT / 16;
  const int ib = i0 + wr * WT, jb = j0 + wc * WT;
  char* base = reinterpret_cast<char*>(A + (int64_t)ib * lda + jb);
  const unsigned lofs =
      (unsigned)((Mfma<T>::lane_row(lane) * lda + (lane & 15)) * (int64_t)sizeof(T));
#pragma unroll
  for (int ti = 0; ti < TI; ++ti)
#pragma unroll
    for (int tj = 0; tj < TI; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        T* p = reinterpret_cast<T*>(
            base + ((int64_t)(ti * 16 + Mfma<T>::v_row(v)) * lda + tj * 16) * (int64_t)sizeof(T) +
            lofs);
        bool ok = true;
        if (!FULL) {
          const int i = ib + ti * 16 + Mfma<T>::row(lane, v);
          const int jj = jb + tj * 16 + (lane & 15);
          ok = i < n && jj < n && jj >= i;
        }
        if (store) {
          if (ok) *p = acc[ti][tj][v];
        } else {
          acc[ti][tj][v] = ok ? *p : T(0);
        }
      }
}

// Tile (I, J) of the trailing triangle: C(i, j) = C[i ldc + j] and
// U(s + q, i) = U[(s + q) ldu + i] in the matrix's own indices (the
// factorization passes A for both; a copy of part of the matrix passes a
// base shifted by the copy's origin), W(q, j) = Wg[q ldw + j].  Every
// element takes the same MFMA chain whichever tile or copy holds it, so any
// tiling of the update gives bitwise the same result.
template <typename T, int TM, int KC>
__device__ __forceinline__ void update_tile_at(T* __restrict__ A, int64_t lda,
                                               const T* __restrict__ U, int64_t ldu,
                                               const T* __restrict__ Wg, int64_t ldw, int s,
                                               int bw, int n, int I, int J,
                                               T (*Us)[KC * (TM + 4)], T (*Ws)[KC * (TM + 4)]) {
  constexpr int LDS_LD = TM + 4;
  constexpr int PER = KC * TM / 256;  // staged elements per thread per operand
  constexpr int WT = TM / 2;          // wavefront tile
  constexpr int TI = WT / 16;         // 16 x 16 MFMA blocks per wavefront side
  typedef typename Mfma<T>::acc_t acc_t;
  const int e = s + bw;
  const int i0 = e + I * TM, j0 = e + J * TM;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wr = wave >> 1, wc = wave & 1;
  const bool full = I < J && j0 + TM <= n;

  acc_t acc[TI][TI];
  if (full)
    c_tile_io<T, TM, true>(A, lda, n, i0, j0, wr, wc, lane, acc, false);
  else
    c_tile_io<T, TM, false>(A, lda, n, i0, j0, wr, wc, lane, acc, false);

  // staging: wavefront w fetches K rows w, w + 4, ... of a chunk, one
  // 64-wide row per load (row address wave-uniform, lane = column)
  static_assert(TM == 64 || TM == 128, "staging: a wavefront loads 64 columns of one row");
  constexpr int RSTEP = 256 / TM;  // rows per pass of the workgroup
  const int qw = (wave * 64) / TM, cw = (wave * 64) % TM + lane;
  const bool iok = full || i0 + cw < n, jok = full || j0 + cw < n;
  T ru[PER], rw[PER];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int kq = k0 + qw + RSTEP * u;
      ru[u] = (kq < bw && iok) ? -U[(int64_t)(s + kq) * ldu + i0 + cw] : T(0);
      rw[u] = (kq < bw && jok) ? Wg[(int64_t)kq * ldw + j0 + cw] : T(0);
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      Us[buf][(qw + RSTEP * u) * LDS_LD + cw] = ru[u];
      Ws[buf][(qw + RSTEP * u) * LDS_LD + cw] = rw[u];
    }
  };

  const int nk = (bw + KC - 1) / KC;
  fetch(0);
  stash(0);
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    const int buf = kc & 1;
    if (kc + 1 < nk) fetch((kc + 1) * KC);
#pragma unroll
    for (int k4 = 0; k4 < KC; k4 += 4) {
      T a[TI], b[TI];
      const int kr = k4 + (lane >> 4);
#pragma unroll
      for (int t = 0; t < TI; ++t) {
        a[t] = Us[buf][kr * LDS_LD + wr * WT + t * 16 + (lane & 15)];
        b[t] = Ws[buf][kr * LDS_LD + wc * WT + t * 16 + (lane & 15)];
      }
#pragma unroll
      for (int ti = 0; ti < TI; ++ti)
#pragma unroll
        for (int tj = 0; tj < TI; ++tj)
          acc[ti][tj] = Mfma<T>::mma(a[ti], b[tj], acc[ti][tj]);
    }
    if (kc + 1 < nk) stash(buf ^ 1);
    __syncthreads();
  }
  if (full)
    c_tile_io<T, TM, true>(A, lda, n, i0, j0, wr, wc, lane, acc, true);
  else
    c_tile_io<T, TM, false>(A, lda, n, i0, j0, wr, wc, lane, acc, true);
}

template <typename T, int TM, int KC, bool SUPER = true>
__device__ __forceinline__ void update_tile(T* __restrict__ A, int64_t lda,
                                            const T* __restrict__ Wg, int64_t ldw, int s,
                                            int bw, int n, int idx, T (*Us)[KC * (TM + 4)],
                                            T (*Ws)[KC * (TM + 4)]) {
  const int m = n - (s + bw);
  if (m <= 0 || bw <= 0) return;
  const int tiles = (m + TM - 1) / TM;
  int I, J;
  if (SUPER) {
    if (idx >= tiles * (tiles + 1) / 2) return;  // the grid covers a size bound
    supertile_of(idx, tiles, &I, &J);
  } else {
    tri_tile(idx, &I, &J);
    if (J >= tiles) return;
  }
  update_tile_at<T, TM, KC>(A, lda, A, lda, Wg, ldw, s, bw, n, I, J, Us, Ws);
}

}  // namespace bbk

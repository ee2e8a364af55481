// bbk_residual.hip -- reconstruction residuals of compact factors on the
// device (gfx950).
//
// Full:    ||P L D L^T P^T - A||_F / max(||A||_F, 1), the residual the
//          reference forms with ExtractRegularFactorsFromCompact,
//          ReconstructOriginal and RelativeError
//          (src/block_ldlt_decomposition.cc:61-78,746-775;
//          tests/block_ldlt_decomposition_test.cc:131-135), without forming
//          L, D or the product: M = L (D L^T) is accumulated tile by tile
//          with FP64 MFMA from the compact factor and compared in place.
// Sampled: the stress test's SampledRelativeResidual
//          (tests/block_ldlt_large_stress_test.cc:160-244): the entries
//          (L D L^T)(pi^-1 i, pi^-1 j) at the same sample positions
//          (four fixed, then std::mt19937(seed + 170003) pairs), one
//          wavefront per entry.
//
// Both read A's upper triangle (row-major, A(i, j) = A[i lda + j], i <= j)
// as the symmetric matrix it defines, and F as written by bbk_?sytrf.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <utility>
#include <vector>

#include "../../include/bbk_ldlt.h"
#include "bbk_handle.hpp"
#include "bbk_update.hpp"

namespace bbk {

constexpr int kResTile = 64;
constexpr int kResKC = 8;

// L(i, p) of the compact factor (CompactLValue,
// tests/block_ldlt_large_stress_test.cc:160-172): 1 on the diagonal, 0
// above it and at the D entry of a 2x2 pivot (p, p+1), else F(p, i).
__device__ __forceinline__ double l_entry(const double* __restrict__ F, int64_t ldf,
                                          const int* __restrict__ piv, int i, int p) {
  if (i == p) return 1.0;
  if (i < p) return 0.0;
  if (piv[p] == 2 && i == p + 1) return 0.0;
  return F[(int64_t)p * ldf + i];
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Tile (I, J), I <= J, of M = L (D L^T): M(i, j) = sum_p L(i, p) E(p, j),
// E(p, j) = D(p, p) L(j, p) + D(p, p+1) L(j, p+1) (piv[p] = 2)
//                           + D(p-1, p) L(j, p-1) (piv[p] = -2),
// over p < min(i0 + 64, n) (L(i, p) = 0 for p > i).  Both operands are
// formed while staged, from rows p, p-1, p+1 of F (contiguous in i, j).
// Epilogue (STORE false): r = M(i, j) - A(perm i, perm j) for i <= j < n,
// weighted 2 off the diagonal (both triangles); part[2 t] = sum w r^2,
// part[2 t + 1] = sum w A^2, fixed order inside the tile.  STORE true:
// out(perm i, perm j) = out(perm j, perm i) = M(i, j), the reconstruction
// P L D L^T P^T (ReconstructOriginal, src/block_ldlt_decomposition.cc:61-78).
template <bool STORE>
__global__ __launch_bounds__(256) void k_ldlt_tiles(int n, const double* __restrict__ A,
                                                    int64_t lda, const double* __restrict__ F,
                                                    int64_t ldf, const int* __restrict__ perm,
                                                    const int* __restrict__ piv,
                                                    double* __restrict__ part,
                                                    double* __restrict__ out, int64_t ldo) {
  constexpr int TM = kResTile, KC = kResKC, LDS_LD = TM + 4;
  constexpr int WT = TM / 2, TI = WT / 16;
  typedef Mfma<double>::acc_t acc_t;
  __shared__ double Ls[2][KC * LDS_LD];
  __shared__ double Es[2][KC * LDS_LD];
  __shared__ double red[2][4];
  int I, J;
  tri_tile(blockIdx.x, &I, &J);
  const int i0 = I * TM, j0 = J * TM;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wr = wave >> 1, wc = wave & 1;
  const int kmax = min(i0 + TM, n);
  acc_t acc[TI][TI];
#pragma unroll
  for (int ti = 0; ti < TI; ++ti)
#pragma unroll
    for (int tj = 0; tj < TI; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[ti][tj][v] = 0.0;

  // wavefront w stages K rows w and w + 4 of a chunk, lane = column
  double rl[2], re[2];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int p = k0 + wave + 4 * u;
      const int i = i0 + lane, j = j0 + lane;
      double a = 0.0, e = 0.0;
      if (p < kmax) {
        if (i < n) a = l_entry(F, ldf, piv, i, p);
        if (j < n) {
          const int pv = piv[p];
          e = F[(int64_t)p * ldf + p] * l_entry(F, ldf, piv, j, p);
          if (pv == 2) e += F[(int64_t)p * ldf + p + 1] * l_entry(F, ldf, piv, j, p + 1);
          if (pv == -2) e += F[(int64_t)(p - 1) * ldf + p] * l_entry(F, ldf, piv, j, p - 1);
        }
      }
      rl[u] = a;
      re[u] = e;
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      Ls[buf][(wave + 4 * u) * LDS_LD + lane] = rl[u];
      Es[buf][(wave + 4 * u) * LDS_LD + lane] = re[u];
    }
  };
  const int nk = (kmax + KC - 1) / KC;
  fetch(0);
  stash(0);
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    const int buf = kc & 1;
    if (kc + 1 < nk) fetch((kc + 1) * KC);
#pragma unroll
    for (int k4 = 0; k4 < KC; k4 += 4) {
      double a[TI], b[TI];
      const int kr = k4 + (lane >> 4);
#pragma unroll
      for (int t = 0; t < TI; ++t) {
        a[t] = Ls[buf][kr * LDS_LD + wr * WT + t * 16 + (lane & 15)];
        b[t] = Es[buf][kr * LDS_LD + wc * WT + t * 16 + (lane & 15)];
      }
#pragma unroll
      for (int ti = 0; ti < TI; ++ti)
#pragma unroll
        for (int tj = 0; tj < TI; ++tj) acc[ti][tj] = Mfma<double>::mma(a[ti], b[tj], acc[ti][tj]);
    }
    if (kc + 1 < nk) stash(buf ^ 1);
    __syncthreads();
  }
  if constexpr (STORE) {
#pragma unroll
    for (int ti = 0; ti < TI; ++ti)
#pragma unroll
      for (int tj = 0; tj < TI; ++tj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = i0 + wr * WT + ti * 16 + Mfma<double>::row(lane, v);
          const int j = j0 + wc * WT + tj * 16 + (lane & 15);
          if (i < n && j < n && i <= j) {
            const int pi = perm[i], pj = perm[j];
            out[(int64_t)pi * ldo + pj] = acc[ti][tj][v];
            out[(int64_t)pj * ldo + pi] = acc[ti][tj][v];
          }
        }
    return;
  }
  double sr = 0.0, sa = 0.0;
#pragma unroll
  for (int ti = 0; ti < TI; ++ti)
#pragma unroll
    for (int tj = 0; tj < TI; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wr * WT + ti * 16 + Mfma<double>::row(lane, v);
        const int j = j0 + wc * WT + tj * 16 + (lane & 15);
        if (i < n && j < n && i <= j) {
          const int pi = perm[i], pj = perm[j];
          const double a = A[(int64_t)min(pi, pj) * lda + max(pi, pj)];
          const double r = acc[ti][tj][v] - a;
          const double w = i == j ? 1.0 : 2.0;
          sr += w * (r * r);
          sa += w * (a * a);
        }
      }
  sr = wave_sum(sr);
  sa = wave_sum(sa);
  if (lane == 0) {
    red[0][wave] = sr;
    red[1][wave] = sa;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * (int64_t)blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    part[2 * (int64_t)blockIdx.x + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// Sum of the per-tile (or per-sample) pairs in a fixed order: thread t
// takes pairs t, t + 256, ...; the 256 partial sums are added pairwise.
// out[0] = sqrt(sum r) / max(sqrt(sum a), 1).
__global__ __launch_bounds__(256) void k_residual_reduce(const double* __restrict__ part,
                                                          int64_t count, double* __restrict__ out) {
  __shared__ double sr[256], sa[256];
  double r = 0.0, a = 0.0;
  for (int64_t t = threadIdx.x; t < count; t += 256) {
    r += part[2 * t];
    a += part[2 * t + 1];
  }
  sr[threadIdx.x] = r;
  sa[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      sa[threadIdx.x] += sa[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sqrt(sr[0]) / fmax(sqrt(sa[0]), 1.0);
}

// The samples' sums in sample order, as the reference accumulates them
// (tests/block_ldlt_large_stress_test.cc:236-243): one thread.
__global__ void k_sampled_reduce(const double* __restrict__ part, int count,
                                 double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double r = 0.0, a = 0.0;
  for (int t = 0; t < count; ++t) {
    r += part[2 * t];
    a += part[2 * t + 1];
  }
  out[0] = sqrt(r) / fmax(sqrt(a), 1.0);
}

// ReconstructedPermutedEntry (tests/block_ldlt_large_stress_test.cc:174-194)
// for sample t, one wavefront: lanes take pivots p = lane, lane + 64, ...
// up to min(row, col) (no later pivot contributes).  out[2 t] = (entry -
// A(i, j))^2, out[2 t + 1] = A(i, j)^2.
__global__ __launch_bounds__(256) void k_sampled_entries(
    int n, const double* __restrict__ A, int64_t lda, const double* __restrict__ F, int64_t ldf,
    const int* __restrict__ perm, const int* __restrict__ piv, const int* __restrict__ samples,
    int count, double* __restrict__ out) {
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= count) return;
  const int orow = samples[2 * t], ocol = samples[2 * t + 1];
  // inverse permutation entries by search: perm is a bijection, so exactly
  // one lane position holds each value
  int frow = -1, fcol = -1;
  for (int i = lane; i < n; i += 64) {
    const int v = perm[i];
    if (v == orow) frow = i;
    if (v == ocol) fcol = i;
  }
  for (int o = 32; o > 0; o >>= 1) {
    frow = max(frow, __shfl_xor(frow, o));
    fcol = max(fcol, __shfl_xor(fcol, o));
  }
  const int pmax = min(frow, fcol);
  double v = 0.0;
  for (int p = lane; p <= pmax; p += 64) {
    const double rl = l_entry(F, ldf, piv, frow, p);
    const double cl = l_entry(F, ldf, piv, fcol, p);
    v += rl * F[(int64_t)p * ldf + p] * cl;
    if (piv[p] == 2) {
      const double off = F[(int64_t)p * ldf + p + 1];
      const double rn = l_entry(F, ldf, piv, frow, p + 1);
      const double cn = l_entry(F, ldf, piv, fcol, p + 1);
      v += off * (rl * cn + rn * cl);
    }
  }
  v = wave_sum(v);
  if (lane == 0) {
    const double a = A[(int64_t)min(orow, ocol) * lda + max(orow, ocol)];
    const double d = v - a;
    out[2 * t] = d * d;
    out[2 * t + 1] = a * a;
  }
}

// ResidualSamples (tests/block_ldlt_large_stress_test.cc:212-227): the
// corners (0, 0), (0, n-1), (n/2, n/2), (n-1, n-1), then pairs drawn with
// std::mt19937(seed + 170003) and std::uniform_int_distribution(0, n-1),
// row first, until `count` samples.
static std::vector<int> residual_samples(int n, int seed, int count) {
  std::vector<std::pair<int, int>> s;
  s.push_back({0, 0});
  s.push_back({0, n - 1});
  s.push_back({n / 2, n / 2});
  s.push_back({n - 1, n - 1});
  std::mt19937 gen(seed + 170003);
  std::uniform_int_distribution<int> dist(0, n - 1);
  while (static_cast<int>(s.size()) < count) {
    const int r = dist(gen);
    const int c = dist(gen);
    s.push_back({r, c});
  }
  std::vector<int> flat;
  flat.reserve(2 * s.size());
  for (const auto& rc : s) {
    flat.push_back(rc.first);
    flat.push_back(rc.second);
  }
  return flat;
}

static int check_residual_args(int n, const double* A, int64_t lda, const double* F, int64_t ldf,
                               const int* perm, const int* piv, const double* out) {
  if (n <= 0) return -2;
  if (!A) return -3;
  if (lda < n) return -4;
  if (!F) return -5;
  if (ldf < n) return -6;
  if (!perm) return -7;
  if (!piv) return -8;
  if (!out) return -9;
  return BBK_OK;
}

static int read_result(Handle* h, const double* d, double* out) {
  hipError_t e = hipMemcpyAsync(out, d, sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return h->hip_fail(e, "read residual");
  return BBK_OK;
}

}  // namespace bbk

using bbk::Handle;

extern "C" {

int bbk_dresidual(bbk_handle_t handle, int n, const double* d_A, int64_t lda,
                  const double* d_F, int64_t ldf, const int* d_perm, const int* d_piv,
                  double* residual) {
  if (!handle) return -1;
  int rc = bbk::check_residual_args(n, d_A, lda, d_F, ldf, d_perm, d_piv, residual);
  if (rc != BBK_OK) return rc;
  Handle* h = reinterpret_cast<Handle*>(handle);
  const int64_t tiles = (n + bbk::kResTile - 1) / bbk::kResTile;
  const int64_t ntile = tiles * (tiles + 1) / 2;
  rc = h->reserve_berr((size_t)(2 * ntile + 1) * sizeof(double));
  if (rc != BBK_OK) return rc;
  double* part = reinterpret_cast<double*>(h->berr_ws);
  double* res = part + 2 * ntile;
  hipLaunchKernelGGL(bbk::k_ldlt_tiles<false>, dim3(static_cast<unsigned>(ntile)), dim3(256), 0,
                     h->stream, n, d_A, lda, d_F, ldf, d_perm, d_piv, part, nullptr, 0);
  hipLaunchKernelGGL(bbk::k_residual_reduce, dim3(1), dim3(256), 0, h->stream, part, ntile, res);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return h->hip_fail(e, "residual kernels");
  return bbk::read_result(h, res, residual);
}

int bbk_dreconstruct(bbk_handle_t handle, int n, const double* d_F, int64_t ldf,
                     const int* d_perm, const int* d_piv, double* d_out, int64_t ldo) {
  if (!handle) return -1;
  if (n <= 0) return -2;
  if (!d_F) return -3;
  if (ldf < n) return -4;
  if (!d_perm) return -5;
  if (!d_piv) return -6;
  if (!d_out) return -7;
  if (ldo < n) return -8;
  Handle* h = reinterpret_cast<Handle*>(handle);
  const int64_t tiles = (n + bbk::kResTile - 1) / bbk::kResTile;
  const int64_t ntile = tiles * (tiles + 1) / 2;
  hipLaunchKernelGGL(bbk::k_ldlt_tiles<true>, dim3(static_cast<unsigned>(ntile)), dim3(256), 0,
                     h->stream, n, nullptr, 0, d_F, ldf, d_perm, d_piv, nullptr, d_out, ldo);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return h->hip_fail(e, "reconstruction kernel");
  return BBK_OK;
}

int bbk_dreconstruct_host(bbk_handle_t handle, int n, const double* F, int64_t ldf,
                          const int* perm, const int* piv, double* out, int64_t ldo) {
  if (!handle) return -1;
  if (n <= 0) return -2;
  if (!F) return -3;
  if (ldf < n) return -4;
  if (!perm) return -5;
  if (!piv) return -6;
  if (!out) return -7;
  if (ldo < n) return -8;
  Handle* h = reinterpret_cast<Handle*>(handle);
  const size_t mat = sizeof(double) * static_cast<size_t>(n) * n;
  double *dF = nullptr, *dO = nullptr;
  int *dp = nullptr, *dv = nullptr;
  hipError_t e = hipMalloc(&dF, mat);
  if (e == hipSuccess) e = hipMalloc(&dO, mat);
  if (e == hipSuccess) e = hipMalloc(&dp, sizeof(int) * n);
  if (e == hipSuccess) e = hipMalloc(&dv, sizeof(int) * n);
  int rc = BBK_OK;
  if (e == hipSuccess)
    e = hipMemcpy2DAsync(dF, sizeof(double) * n, F, sizeof(double) * ldf, sizeof(double) * n, n,
                         hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dp, perm, sizeof(int) * n, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dv, piv, sizeof(int) * n, hipMemcpyHostToDevice, h->stream);
  if (e != hipSuccess) {
    rc = h->hip_fail(e, "reconstruction upload");
  } else {
    rc = bbk_dreconstruct(handle, n, dF, n, dp, dv, dO, n);
    if (rc == BBK_OK) {
      e = hipMemcpy2DAsync(out, sizeof(double) * ldo, dO, sizeof(double) * n,
                           sizeof(double) * n, n, hipMemcpyDeviceToHost, h->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
      if (e != hipSuccess) rc = h->hip_fail(e, "reconstruction download");
    }
  }
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(dF);
  (void)hipFree(dO);
  (void)hipFree(dp);
  (void)hipFree(dv);
  return rc;
}

int bbk_residual_samples(int n, int seed, int count, int* rows, int* cols) {
  if (n <= 0) return -1;
  if (count < 4) return -3;
  if (!rows) return -4;
  if (!cols) return -5;
  const std::vector<int> smp = bbk::residual_samples(n, seed, count);
  for (int t = 0; t < count; ++t) {
    rows[t] = smp[2 * t];
    cols[t] = smp[2 * t + 1];
  }
  return BBK_OK;
}

int bbk_dsampled_residual(bbk_handle_t handle, int n, const double* d_A, int64_t lda,
                          const double* d_F, int64_t ldf, const int* d_perm, const int* d_piv,
                          int seed, int count, double* residual) {
  if (!handle) return -1;
  int rc = bbk::check_residual_args(n, d_A, lda, d_F, ldf, d_perm, d_piv, residual);
  if (rc != BBK_OK) return rc;
  if (count < 4) return -11;
  Handle* h = reinterpret_cast<Handle*>(handle);
  const std::vector<int> smp = bbk::residual_samples(n, seed, count);
  const size_t sbytes = (smp.size() * sizeof(int) + 255) / 256 * 256;
  rc = h->reserve_berr(sbytes + (size_t)(2 * count + 1) * sizeof(double));
  if (rc != BBK_OK) return rc;
  int* d_s = reinterpret_cast<int*>(h->berr_ws);
  double* part = reinterpret_cast<double*>(reinterpret_cast<char*>(h->berr_ws) + sbytes);
  double* res = part + 2 * (int64_t)count;
  hipError_t e = hipMemcpyAsync(d_s, smp.data(), smp.size() * sizeof(int), hipMemcpyHostToDevice,
                                h->stream);
  if (e != hipSuccess) return h->hip_fail(e, "upload samples");
  hipLaunchKernelGGL(bbk::k_sampled_entries, dim3((count + 3) / 4), dim3(256), 0, h->stream, n,
                     d_A, lda, d_F, ldf, d_perm, d_piv, d_s, count, part);
  hipLaunchKernelGGL(bbk::k_sampled_reduce, dim3(1), dim3(64), 0, h->stream, part, count, res);
  e = hipGetLastError();
  if (e != hipSuccess) return h->hip_fail(e, "sampled residual kernels");
  // the upload came from pageable memory: wait before smp goes away
  return bbk::read_result(h, res, residual);
}

}  // extern "C"

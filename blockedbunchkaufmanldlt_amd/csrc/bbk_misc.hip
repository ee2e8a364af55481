// bbk_misc.hip -- kernels.h counterparts (src/kernels.cc) and the
// synthetic-input generator.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cstdint>

#include "../../include/bbk_ldlt.h"
#include "bbk_handle.hpp"
#include "bbk_kernels.hpp"
#include "bbk_update.hpp"

namespace bbk {

// kernels::ScaleOneByOnePivot (non-MKL path), src/kernels.cc:11-31:
// exact-zero pivot -> zeros, else true division.
__global__ void k_scale_1x1(int count, double pivot, const double* src,
                            int ss, double* dst, int ds) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  dst[(int64_t)i * ds] = (fabs(pivot) == 0.0) ? 0.0 : src[(int64_t)i * ss] / pivot;
}

// kernels::ScaleTwoByTwoPivot, src/kernels.cc:33-65 (its own formula,
// (v1*c - v2*b)/det, not the factorization's inverse form).
__global__ void k_scale_2x2(int count, double a, double b, double c,
                            const double* s1, int s1s, const double* s2,
                            int s2s, double* d1, int d1s, double* d2,
                            int d2s) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const double det = a * c - b * b;
  if (fabs(det) == 0.0) {
    d1[(int64_t)i * d1s] = 0.0;
    d2[(int64_t)i * d2s] = 0.0;
    return;
  }
  const double v1 = s1[(int64_t)i * s1s];
  const double v2 = s2[(int64_t)i * s2s];
  d1[(int64_t)i * d1s] = (v1 * c - v2 * b) / det;
  d2[(int64_t)i * d2s] = (-v1 * b + v2 * a) / det;
}

// kernels::UpdateTrailingBlock, src/kernels.cc:67-91:
// target(i,j) -= sum_k left(k,i) * work(k,j).  One 64 x 64 target tile per
// workgroup of four wavefronts (32 x 32 each, 2 x 2 blocks of the binary64
// 16x16x4 MFMA).  The accumulators start from the target tile and the left
// operand is negated while it is staged, so the MFMA chain computes
// target - left^T work in place; K is staged through two LDS buffers in
// chunks of 8 rows, the next chunk's global loads overlapping the current
// chunk's MFMAs.  Every access is guarded against rows, cols and inner.
__global__ __launch_bounds__(256) void k_update_trailing(
    int rows, int cols, int inner, const double* __restrict__ left, int ls,
    const double* __restrict__ work, int ws, double* __restrict__ target, int ts) {
  constexpr int TM = 64, KC = 8, LD = TM + 4, WT = TM / 2, TI = WT / 16;
  constexpr int PER = KC * TM / 256;  // staged elements per thread per operand
  constexpr int RSTEP = 256 / TM;     // K rows per pass of the workgroup
  typedef Mfma<double>::acc_t acc_t;
  __shared__ __align__(16) double Ls[2][KC * LD];
  __shared__ __align__(16) double Rs[2][KC * LD];
  const int i0 = blockIdx.y * TM, j0 = blockIdx.x * TM;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wr = wave >> 1, wc = wave & 1;

  acc_t acc[TI][TI];
#pragma unroll
  for (int ti = 0; ti < TI; ++ti)
#pragma unroll
    for (int tj = 0; tj < TI; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wr * WT + ti * 16 + Mfma<double>::row(lane, v);
        const int j = j0 + wc * WT + tj * 16 + (lane & 15);
        acc[ti][tj][v] = (i < rows && j < cols) ? target[(int64_t)i * ts + j] : 0.0;
      }

  // wavefront w stages K rows w, w + 4 of a chunk, lane = column
  const int qw = wave, cw = lane;
  const bool iok = i0 + cw < rows, jok = j0 + cw < cols;
  double ru[PER], rw[PER];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int kq = k0 + qw + RSTEP * u;
      ru[u] = (kq < inner && iok) ? -left[(int64_t)kq * ls + i0 + cw] : 0.0;
      rw[u] = (kq < inner && jok) ? work[(int64_t)kq * ws + j0 + cw] : 0.0;
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      Ls[buf][(qw + RSTEP * u) * LD + cw] = ru[u];
      Rs[buf][(qw + RSTEP * u) * LD + cw] = rw[u];
    }
  };
  const int nk = (inner + KC - 1) / KC;
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
        a[t] = Ls[buf][kr * LD + wr * WT + t * 16 + (lane & 15)];
        b[t] = Rs[buf][kr * LD + wc * WT + t * 16 + (lane & 15)];
      }
#pragma unroll
      for (int ti = 0; ti < TI; ++ti)
#pragma unroll
        for (int tj = 0; tj < TI; ++tj) acc[ti][tj] = Mfma<double>::mma(a[ti], b[tj], acc[ti][tj]);
    }
    if (kc + 1 < nk) stash(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int ti = 0; ti < TI; ++ti)
#pragma unroll
    for (int tj = 0; tj < TI; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wr * WT + ti * 16 + Mfma<double>::row(lane, v);
        const int j = j0 + wc * WT + tj * 16 + (lane & 15);
        if (i < rows && j < cols) target[(int64_t)i * ts + j] = acc[ti][tj][v];
      }
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// A(i,j) = A(j,i) = scale * u(seed, min(i,j), max(i,j)),
// u = 2 * (key >> 11) * 2^-53 - 1, key = splitmix64(splitmix64(seed) ^ (i << 32 | j)).
__global__ void k_gen_synth(int n, uint64_t seed, double scale, int zstride,
                            double* A, int64_t lda) {
  // grid-stride: n^2 passes the 2^32 work-items of one launch at n = 65536
  const int64_t total = (int64_t)n * n;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += step) {
    const int i = static_cast<int>(idx / n), j = static_cast<int>(idx % n);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const uint64_t key = splitmix64(splitmix64(seed) ^
                                    ((uint64_t)lo << 32 | (uint32_t)hi));
    double v = scale * (2.0 * (static_cast<double>(key >> 11) * 0x1.0p-53) - 1.0);
    if (i == j && zstride > 0 && i % zstride == 0) v = 0.0;
    A[(int64_t)i * lda + j] = v;
  }
}

}  // namespace bbk

using bbk::Handle;

extern "C" {

int bbk_kernel_scale_1x1(bbk_handle_t handle, int count, double pivot,
                         const double* d_src, int src_stride, double* d_dst,
                         int dst_stride) {
  if (!handle) return -1;
  if (count <= 0) return BBK_OK;  // no-op, as src/kernels.cc:14-16
  Handle* h = reinterpret_cast<Handle*>(handle);
  hipLaunchKernelGGL(bbk::k_scale_1x1, dim3((count + 255) / 256), dim3(256), 0,
                     h->stream, count, pivot, d_src, src_stride, d_dst,
                     dst_stride);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? BBK_OK : h->hip_fail(e, "scale_1x1");
}

int bbk_kernel_scale_2x2(bbk_handle_t handle, int count, double a, double b,
                         double c, const double* d_s1, int s1_stride,
                         const double* d_s2, int s2_stride, double* d_d1,
                         int d1_stride, double* d_d2, int d2_stride) {
  if (!handle) return -1;
  if (count <= 0) return BBK_OK;  // src/kernels.cc:41-43
  Handle* h = reinterpret_cast<Handle*>(handle);
  hipLaunchKernelGGL(bbk::k_scale_2x2, dim3((count + 255) / 256), dim3(256), 0,
                     h->stream, count, a, b, c, d_s1, s1_stride, d_s2,
                     s2_stride, d_d1, d1_stride, d_d2, d2_stride);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? BBK_OK : h->hip_fail(e, "scale_2x2");
}

int bbk_kernel_update_trailing(bbk_handle_t handle, int rows, int cols,
                               int inner, const double* d_left,
                               int left_stride, const double* d_work,
                               int work_stride, double* d_target,
                               int target_stride) {
  if (!handle) return -1;
  if (rows <= 0 || cols <= 0 || inner <= 0) return BBK_OK;  // :72-74
  Handle* h = reinterpret_cast<Handle*>(handle);
  if ((rows + 63) / 64 > 65535) return h->fail(BBK_ERR_UNSUPPORTED, "update_trailing: rows > 4194240");
  dim3 grid((cols + 63) / 64, (rows + 63) / 64);
  hipLaunchKernelGGL(bbk::k_update_trailing, grid, dim3(256), 0, h->stream,
                     rows, cols, inner, d_left, left_stride, d_work,
                     work_stride, d_target, target_stride);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? BBK_OK : h->hip_fail(e, "update_trailing");
}

int bbk_dgen_synth(bbk_handle_t handle, int n, uint64_t seed, double scale,
                   int zero_diag_stride, double* d_A, int64_t lda) {
  if (!handle) return -1;
  if (n <= 0) return -2;
  if (!d_A) return -6;
  if (lda < n) return -7;
  Handle* h = reinterpret_cast<Handle*>(handle);
  const int64_t total = (int64_t)n * n;
  const int64_t blocks = std::min<int64_t>((total + 255) / 256, int64_t(1) << 20);
  hipLaunchKernelGGL(bbk::k_gen_synth, dim3((unsigned)blocks), dim3(256), 0, h->stream, n, seed,
                     scale, zero_diag_stride, d_A, lda);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? BBK_OK : h->hip_fail(e, "gen_synth");
}

}  // extern "C"

// ---- host-pointer forms of the kernels.h helpers -----------------------
namespace {

// Device mirror of a strided host vector of `count` elements.
struct Mirror {
  double* d = nullptr;
  size_t elems = 0;
  hipError_t err = hipSuccess;
  Mirror(const double* h, int count, int stride, hipStream_t st) {
    elems = count <= 0 ? 0 : static_cast<size_t>(count - 1) * static_cast<size_t>(stride) + 1;
    if (elems == 0) return;
    err = hipMalloc(&d, elems * sizeof(double));
    if (err == hipSuccess)
      err = hipMemcpyAsync(d, h, elems * sizeof(double), hipMemcpyHostToDevice, st);
  }
  hipError_t back(double* h, hipStream_t st) {
    if (elems == 0) return hipSuccess;
    hipError_t e = hipMemcpyAsync(h, d, elems * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
  }
  ~Mirror() {
    if (d) (void)hipFree(d);
  }
};

// extent of a row-major operand whose rows are `stride` apart
size_t extent_rows(int nrows, int ncols, int stride) {
  return static_cast<size_t>(nrows - 1) * static_cast<size_t>(stride) + ncols;
}

}  // namespace

extern "C" {

int bbk_kernel_scale_1x1_host(bbk_handle_t handle, int count, double pivot, const double* src,
                              int src_stride, double* dst, int dst_stride) {
  if (!handle) return -1;
  if (count <= 0) return BBK_OK;
  if (!src) return -4;
  if (src_stride < 0) return -5;
  if (!dst) return -6;
  if (dst_stride < 0) return -7;
  Handle* h = reinterpret_cast<Handle*>(handle);
  Mirror s(src, count, src_stride, h->stream), d(dst, count, dst_stride, h->stream);
  if (s.err != hipSuccess || d.err != hipSuccess) return h->hip_fail(s.err != hipSuccess ? s.err : d.err, "device copy");
  int rc = bbk_kernel_scale_1x1(handle, count, pivot, s.d, src_stride, d.d, dst_stride);
  if (rc != BBK_OK) return rc;
  hipError_t e = d.back(dst, h->stream);
  return e == hipSuccess ? BBK_OK : h->hip_fail(e, "copy back");
}

int bbk_kernel_scale_2x2_host(bbk_handle_t handle, int count, double a, double b, double c,
                              const double* s1, int s1_stride, const double* s2, int s2_stride,
                              double* d1, int d1_stride, double* d2, int d2_stride) {
  if (!handle) return -1;
  if (count <= 0) return BBK_OK;
  if (!s1) return -6;
  if (s1_stride < 0) return -7;
  if (!s2) return -8;
  if (s2_stride < 0) return -9;
  if (!d1) return -10;
  if (d1_stride < 0) return -11;
  if (!d2) return -12;
  if (d2_stride < 0) return -13;
  Handle* h = reinterpret_cast<Handle*>(handle);
  Mirror m1(s1, count, s1_stride, h->stream), m2(s2, count, s2_stride, h->stream);
  Mirror o1(d1, count, d1_stride, h->stream), o2(d2, count, d2_stride, h->stream);
  for (hipError_t e : {m1.err, m2.err, o1.err, o2.err})
    if (e != hipSuccess) return h->hip_fail(e, "device copy");
  int rc = bbk_kernel_scale_2x2(handle, count, a, b, c, m1.d, s1_stride, m2.d, s2_stride, o1.d,
                                d1_stride, o2.d, d2_stride);
  if (rc != BBK_OK) return rc;
  hipError_t e = o1.back(d1, h->stream);
  if (e == hipSuccess) e = o2.back(d2, h->stream);
  return e == hipSuccess ? BBK_OK : h->hip_fail(e, "copy back");
}

int bbk_kernel_update_trailing_host(bbk_handle_t handle, int rows, int cols, int inner,
                                    const double* left, int left_stride, const double* work,
                                    int work_stride, double* target, int target_stride) {
  if (!handle) return -1;
  if (rows <= 0 || cols <= 0 || inner <= 0) return BBK_OK;
  if (!left) return -5;
  if (left_stride < 0) return -6;
  if (!work) return -7;
  if (work_stride < 0) return -8;
  if (!target) return -9;
  if (target_stride < 0) return -10;
  Handle* h = reinterpret_cast<Handle*>(handle);
  const size_t el = extent_rows(inner, rows, left_stride);
  const size_t ew = extent_rows(inner, cols, work_stride);
  const size_t et = extent_rows(rows, cols, target_stride);
  double *dl = nullptr, *dw = nullptr, *dt = nullptr;
  hipError_t e = hipMalloc(&dl, el * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dw, ew * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dt, et * sizeof(double));
  if (e == hipSuccess)
    e = hipMemcpyAsync(dl, left, el * sizeof(double), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dw, work, ew * sizeof(double), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dt, target, et * sizeof(double), hipMemcpyHostToDevice, h->stream);
  int rc = BBK_OK;
  if (e != hipSuccess) {
    rc = h->hip_fail(e, "device copy");
  } else {
    rc = bbk_kernel_update_trailing(handle, rows, cols, inner, dl, left_stride, dw, work_stride,
                                    dt, target_stride);
    if (rc == BBK_OK) {
      e = hipMemcpyAsync(target, dt, et * sizeof(double), hipMemcpyDeviceToHost, h->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
      if (e != hipSuccess) rc = h->hip_fail(e, "copy back");
    }
  }
  (void)hipStreamSynchronize(h->stream);
  if (dl) (void)hipFree(dl);
  if (dw) (void)hipFree(dw);
  if (dt) (void)hipFree(dt);
  return rc;
}

}  // extern "C"

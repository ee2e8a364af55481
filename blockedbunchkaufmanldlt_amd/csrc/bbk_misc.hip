// bbk_misc.hip -- kernels.h counterparts (src/kernels.cc) and the
// synthetic-input generator.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cstdint>

#include "../../include/bbk_ldlt.h"
#include "bbk_handle.hpp"
#include "bbk_kernels.hpp"

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
// target(i,j) -= sum_k left(k,i) * work(k,j); 16x16 output per wave via
// the f64 MFMA (16x16x4), one wave per 16x16 tile.
__global__ __launch_bounds__(64) void k_update_trailing(
    int rows, int cols, int inner, const double* left, int ls,
    const double* work, int ws, double* target, int ts) {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  const int i0 = blockIdx.y * 16, j0 = blockIdx.x * 16;
  const int lane = threadIdx.x;
  acc_t acc = {0, 0, 0, 0};
  for (int k0 = 0; k0 < inner; k0 += 4) {
    const int k = k0 + (lane >> 4);
    const int ii = i0 + (lane & 15), jj = j0 + (lane & 15);
    const double a = (k < inner && ii < rows) ? left[(int64_t)k * ls + ii] : 0.0;
    const double b = (k < inner && jj < cols) ? work[(int64_t)k * ws + jj] : 0.0;
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  }
  for (int v = 0; v < 4; ++v) {
    // f64 16x16x4 output: element v of a lane is row (lane / 16) + 4 v
    const int i = i0 + (lane >> 4) + 4 * v, j = j0 + (lane & 15);
    if (i < rows && j < cols) target[(int64_t)i * ts + j] -= acc[v];
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
  dim3 grid((cols + 15) / 16, (rows + 15) / 16);
  hipLaunchKernelGGL(bbk::k_update_trailing, grid, dim3(64), 0, h->stream,
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

// bbk_peaks -- achievable peaks of the GPU it runs on, for the rooflines the
// benchmark reports beside the datasheet peaks:
//   hbm_copy   a float4 copy between two 2 GiB buffers (far beyond every
//              cache), bytes read + written per second;
//   fp64_mfma  v_mfma_f64_16x16x4f64, independent accumulator chains per
//              wavefront;
//   fp32_mfma  the faster of v_mfma_f32_16x16x4f32 and
//              v_mfma_f32_32x32x2f32;
//   fp64_valu  v_fma_f64, independent chains per lane (the vector unit's
//              binary64 rate, which the panel's row products use).
// Each over a few launch shapes (loads in flight, workgroups per CU,
// chains per wavefront), the best of five timed launches of each (HIP
// events, after one untimed launch); the best shape is reported.
// Prints one JSON line.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

namespace {

#define CHECK(x)                                                      \
  do {                                                                \
    if ((x) != hipSuccess) {                                          \
      std::fprintf(stderr, "HIP call failed at line %d\n", __LINE__); \
      return 1;                                                       \
    }                                                                 \
  } while (0)

// U float4 loads of a thread in flight before its stores; NT: the stores
// non-temporal.
typedef float v4f_t __attribute__((ext_vector_type(4)));

template <int U, bool NT>
__global__ __launch_bounds__(256) void k_copy(const v4f_t* __restrict__ a,
                                              v4f_t* __restrict__ b, size_t count) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < count; i += U * stride) {
    v4f_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = a[i + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (NT)
        __builtin_nontemporal_store(v[u], &b[i + u * stride]);
      else
        b[i + u * stride] = v[u];
    }
  }
  for (; i < count; i += stride) b[i] = a[i];
}

// A workgroup copies U * 256 consecutive elements per step (its U loads
// per thread adjacent in memory), steps strided over the grid.
template <int U>
__global__ __launch_bounds__(256) void k_copy_blocked(const v4f_t* __restrict__ a,
                                                      v4f_t* __restrict__ b, size_t count) {
  const size_t step = (size_t)U * 256;
  for (size_t base = (size_t)blockIdx.x * step; base + step <= count;
       base += (size_t)gridDim.x * step) {
    v4f_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = a[base + (size_t)u * 256 + threadIdx.x];
#pragma unroll
    for (int u = 0; u < U; ++u) b[base + (size_t)u * 256 + threadIdx.x] = v[u];
  }
}

// CH independent accumulator chains per wavefront.
template <int CH>
__global__ __launch_bounds__(256) void k_mfma_f64(int iters, double* out) {
  typedef double v4d __attribute__((ext_vector_type(4)));
  v4d acc[CH];
  const double a = 1e-3 * threadIdx.x, b = 1.0 + 1e-6 * blockIdx.x;
#pragma unroll
  for (int c = 0; c < CH; ++c) acc[c] = v4d{0.0, 0.0, 0.0, 0.0};
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < CH; ++c)
      acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < CH; ++c) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// VF independent v_fma_f64 chains per lane.
template <int VF>
__global__ __launch_bounds__(256) void k_valu_f64(int iters, double* out) {
  double x[VF];
  const double a = 1.0 + 1e-9 * threadIdx.x, b = 1e-9 * blockIdx.x;
#pragma unroll
  for (int c = 0; c < VF; ++c) x[c] = 1e-3 * c;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < VF; ++c) x[c] = __builtin_fma(x[c], a, b);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < VF; ++c) s += x[c];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <int CH>
__global__ __launch_bounds__(256) void k_mfma_f32_16(int iters, float* out) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  v4f acc[CH];
  const float a = 1e-3f * threadIdx.x, b = 1.0f + 1e-6f * blockIdx.x;
#pragma unroll
  for (int c = 0; c < CH; ++c) acc[c] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < CH; ++c)
      acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[c], 0, 0, 0);
  }
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < CH; ++c) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <int CH>
__global__ __launch_bounds__(256) void k_mfma_f32_32(int iters, float* out) {
  typedef float v16f __attribute__((ext_vector_type(16)));
  v16f acc[CH];
  const float a = 1e-3f * threadIdx.x, b = 1.0f + 1e-6f * blockIdx.x;
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[c][v] = 0.f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < CH; ++c)
      acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[c], 0, 0, 0);
  }
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int v = 0; v < 16; ++v) s += acc[c][v];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Best of five timed launches of `launch` (after one untimed), in ms.
template <typename F>
int best_ms(F launch, float* ms) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  launch();
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  float best = 1e30f;
  for (int r = 0; r < 5; ++r) {
    CHECK(hipEventRecord(e0, nullptr));
    launch();
    CHECK(hipEventRecord(e1, nullptr));
    CHECK(hipEventSynchronize(e1));
    float t = 0.f;
    CHECK(hipEventElapsedTime(&t, e0, e1));
    best = std::min(best, t);
  }
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  *ms = best;
  return 0;
}

}  // namespace

int main() {
  int dev = 0, cus = 0;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));

  // HBM: copy between two 2 GiB buffers; 1, 4 or 8 loads in flight per
  // thread, grid-strided or workgroup-contiguous, plain or non-temporal
  // stores, 8, 16 or 64 workgroups per CU (count is a multiple of every
  // workgroup step, so the blocked form copies everything)
  const size_t bytes = size_t(2) << 30;
  const size_t count = bytes / sizeof(v4f_t);
  v4f_t *a = nullptr, *b = nullptr;
  CHECK(hipMalloc(&a, bytes));
  CHECK(hipMalloc(&b, bytes));
  CHECK(hipMemset(a, 0, bytes));
  float copy_ms = 1e30f;
  for (int per_cu : {8, 16, 64}) {
    const int grid = cus * per_cu;
    float t[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (best_ms([&] { hipLaunchKernelGGL((k_copy<1, false>), dim3(grid), dim3(256), 0, nullptr,
                                         a, b, count); }, &t[0]) ||
        best_ms([&] { hipLaunchKernelGGL((k_copy<4, false>), dim3(grid), dim3(256), 0, nullptr,
                                         a, b, count); }, &t[1]) ||
        best_ms([&] { hipLaunchKernelGGL((k_copy<8, false>), dim3(grid), dim3(256), 0, nullptr,
                                         a, b, count); }, &t[2]) ||
        best_ms([&] { hipLaunchKernelGGL((k_copy<4, true>), dim3(grid), dim3(256), 0, nullptr,
                                         a, b, count); }, &t[3]) ||
        best_ms([&] { hipLaunchKernelGGL((k_copy<8, true>), dim3(grid), dim3(256), 0, nullptr,
                                         a, b, count); }, &t[4]) ||
        best_ms([&] { hipLaunchKernelGGL((k_copy_blocked<4>), dim3(grid), dim3(256), 0, nullptr,
                                         a, b, count); }, &t[5]) ||
        best_ms([&] { hipLaunchKernelGGL((k_copy_blocked<8>), dim3(grid), dim3(256), 0, nullptr,
                                         a, b, count); }, &t[6]))
      return 1;
    for (float v : t) copy_ms = std::min(copy_ms, v);
  }
  CHECK(hipFree(a));
  CHECK(hipFree(b));

  // MFMA: 4 or 8 workgroups of 4 wavefronts per CU, 4 or 8 accumulator
  // chains per wavefront; the best rate of each instruction
  void* out = nullptr;
  CHECK(hipMalloc(&out, (size_t)cus * 8 * 256 * sizeof(double)));
  double f64 = 0.0, f32a = 0.0, f32b = 0.0, v64 = 0.0;
  for (int per_cu : {4, 8}) {
    const int grid = cus * per_cu, threads = 256;
    const double waves = (double)grid * threads / 64.0;
    // about 1.3e12 (binary64) or 2.7e12 / 5.4e12 (binary32) flop per launch
    const int it64 = 160000 / per_cu, it32 = 320000 / per_cu;
    float t = 0.f;
    if (best_ms([&] { hipLaunchKernelGGL(k_mfma_f64<4>, dim3(grid), dim3(threads), 0, nullptr,
                                         it64, static_cast<double*>(out)); }, &t))
      return 1;
    f64 = std::max(f64, waves * it64 * 4 * 2048.0 / (t * 1e-3) / 1e12);
    if (best_ms([&] { hipLaunchKernelGGL(k_mfma_f64<8>, dim3(grid), dim3(threads), 0, nullptr,
                                         it64 / 2, static_cast<double*>(out)); }, &t))
      return 1;
    f64 = std::max(f64, waves * (it64 / 2) * 8 * 2048.0 / (t * 1e-3) / 1e12);
    if (best_ms([&] { hipLaunchKernelGGL(k_mfma_f32_16<4>, dim3(grid), dim3(threads), 0, nullptr,
                                         it32, static_cast<float*>(out)); }, &t))
      return 1;
    f32a = std::max(f32a, waves * it32 * 4 * 2048.0 / (t * 1e-3) / 1e12);
    if (best_ms([&] { hipLaunchKernelGGL(k_mfma_f32_32<4>, dim3(grid), dim3(threads), 0, nullptr,
                                         it32, static_cast<float*>(out)); }, &t))
      return 1;
    f32b = std::max(f32b, waves * it32 * 4 * 4096.0 / (t * 1e-3) / 1e12);
    // about 1.7e11 flop per launch
    const int itv = 40000 / per_cu;
    if (best_ms([&] { hipLaunchKernelGGL(k_valu_f64<8>, dim3(grid), dim3(threads), 0, nullptr,
                                         itv, static_cast<double*>(out)); }, &t))
      return 1;
    v64 = std::max(v64, (double)grid * threads * itv * 8 * 2.0 / (t * 1e-3) / 1e12);
  }
  CHECK(hipFree(out));

  const double hbm = 2.0 * (double)bytes / (copy_ms * 1e-3) / 1e9;
  std::printf("{\"hbm_copy_GB/s\": %.1f, \"fp64_mfma_TFLOP/s\": %.2f, "
              "\"fp32_mfma_16x16x4_TFLOP/s\": %.2f, \"fp32_mfma_32x32x2_TFLOP/s\": %.2f, "
              "\"fp32_mfma_TFLOP/s\": %.2f, \"fp64_valu_TFLOP/s\": %.2f, \"cus\": %d}\n",
              hbm, f64, f32a, f32b, std::max(f32a, f32b), v64, cus);
  return 0;
}

// bbk_profile_driver -- minimal C driver of the C ABI for profiler runs
// (rocprofv3 --kernel-trace / --pmc): generates a synthetic symmetric
// matrix on the device, factors it `reps` times, and exits cleanly.
//   bbk_profile_driver <n> <reps> <f64|f32> [nb]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "bbk_ldlt.h"

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 4096;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 1;
  const bool f32 = argc > 3 && std::strcmp(argv[3], "f32") == 0;
  const int nb = argc > 4 ? std::atoi(argv[4]) : 0;
  int64_t lda = (n + 63) / 64 * 64;
  if (lda % 1024 == 0) lda += 64;
  bbk_handle_t h;
  if (bbk_create(&h, nullptr) != BBK_OK) return 1;
  double* A0 = nullptr;
  double* A = nullptr;
  float* S = nullptr;
  int *perm = nullptr, *piv = nullptr;
  const size_t bytes = sizeof(double) * (size_t)lda * n;
  if (hipMalloc(&A0, bytes) != hipSuccess || hipMalloc(&A, bytes) != hipSuccess ||
      hipMalloc(&S, bytes / 2) != hipSuccess || hipMalloc(&perm, 4 * n) != hipSuccess ||
      hipMalloc(&piv, 4 * n) != hipSuccess) {
    std::fprintf(stderr, "allocation failed\n");
    return 1;
  }
  bbk_dgen_synth(h, n, 20260, 1.0, 7, A0, lda);
  bbk_synchronize(h);
  int rc = 0;
  for (int r = 0; r < reps && rc == 0; ++r) {
    if (f32) {
      // binary32 copy of the row-major matrix (upper triangle used)
      float* tmp = (float*)std::malloc(sizeof(float) * (size_t)lda * n);
      double* hd = (double*)std::malloc(bytes);
      hipMemcpy(hd, A0, bytes, hipMemcpyDeviceToHost);
      for (size_t i = 0; i < (size_t)lda * n; ++i) tmp[i] = (float)hd[i];
      hipMemcpy(S, tmp, bytes / 2, hipMemcpyHostToDevice);
      std::free(tmp);
      std::free(hd);
      rc = bbk_ssytrf(h, n, nb, S, lda, perm, piv);
    } else {
      hipMemcpy(A, A0, bytes, hipMemcpyDeviceToDevice);
      rc = bbk_dsytrf(h, n, nb, A, lda, perm, piv);
    }
    if (rc == 0) rc = bbk_synchronize(h);
    std::printf("rep %d rc %d\n", r, rc);
    std::fflush(stdout);
  }
  if (rc != 0) std::fprintf(stderr, "error: %s\n", bbk_last_error(h));
  bbk_destroy(h);
  hipFree(A0);
  hipFree(A);
  hipFree(S);
  hipFree(perm);
  hipFree(piv);
  std::printf("done\n");
  return rc;
}

// bbk_profile_driver -- minimal C driver of the C ABI for profiler runs
// (rocprofv3 --kernel-trace / --pmc): generates a synthetic symmetric
// matrix on the device, runs one operation on it `reps` times, and exits
// cleanly.
//   bbk_profile_driver <n> <reps> <f64|f32|mixed|sysv> [nb]
// f64 / f32: factorization (bbk_dsytrf / bbk_ssytrf);
// mixed: bbk_dsysv_mixed, one right-hand side, tolerance 1e-12;
// sysv: bbk_dsysv, the binary64 solve, one right-hand side.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "bbk_ldlt.h"

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 4096;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 1;
  const char* mode = argc > 3 ? argv[3] : "f64";
  const bool f32 = std::strcmp(mode, "f32") == 0;
  const bool mixed = std::strcmp(mode, "mixed") == 0;
  const bool sysv = std::strcmp(mode, "sysv") == 0;
  const int nb = argc > 4 ? std::atoi(argv[4]) : 0;
  int64_t lda = (n + 63) / 64 * 64;
  if (lda % 1024 == 0) lda += 64;
  bbk_handle_t h;
  if (bbk_create(&h, nullptr) != BBK_OK) return 1;
  double* A0 = nullptr;
  double* A = nullptr;
  float* S = nullptr;
  double *b = nullptr, *x = nullptr;
  int *perm = nullptr, *piv = nullptr;
  const size_t bytes = sizeof(double) * (size_t)lda * n;
  if (hipMalloc(&A0, bytes) != hipSuccess || hipMalloc(&A, bytes) != hipSuccess ||
      hipMalloc(&S, bytes / 2) != hipSuccess || hipMalloc(&perm, 4 * n) != hipSuccess ||
      hipMalloc(&piv, 4 * n) != hipSuccess || hipMalloc(&b, 8 * n) != hipSuccess ||
      hipMalloc(&x, 8 * n) != hipSuccess) {
    std::fprintf(stderr, "allocation failed\n");
    return 1;
  }
  bbk_dgen_synth(h, n, 20260, 1.0, 7, A0, lda);
  bbk_synchronize(h);
  {
    double* hb = (double*)std::malloc(8 * (size_t)n);
    for (int i = 0; i < n; ++i) hb[i] = 1.0 + (double)((i * 37) % 101) / 101.0;
    hipMemcpy(b, hb, 8 * (size_t)n, hipMemcpyHostToDevice);
    std::free(hb);
  }
  int rc = 0;
  for (int r = 0; r < reps && rc == 0; ++r) {
    if (mixed || sysv) {
      int iters = 0;
      double be = 0.0;
      rc = mixed ? bbk_dsysv_mixed(h, n, 1, A0, lda, b, 1, x, 1, 1e-12, 30, &iters, &be)
                 : bbk_dsysv(h, n, 1, A0, lda, b, 1, x, 1, &be);
      std::printf("rep %d iters %d berr %.3e\n", r, iters, be);
    } else if (f32) {
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
  hipFree(b);
  hipFree(x);
  std::printf("done\n");
  std::fflush(stdout);
  return rc;
}

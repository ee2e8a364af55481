// bbk_profile_driver -- minimal C driver of the C ABI for profiler runs
// (rocprofv3 --kernel-trace / --pmc): builds synthetic inputs on the
// device, runs one operation `reps` times, and exits cleanly.
//   bbk_profile_driver <n> <reps> <mode> [nb]
// f64 / f32  factorization of one order-n matrix (bbk_dsytrf / bbk_ssytrf);
// mixed      bbk_dsysv_mixed, one right-hand side, tolerance 1e-12;
// sysv       bbk_dsysv, the binary64 solve, one right-hand side;
// c4         the small batch, binary64 then binary32: 4096 matrices with
//            orders in [16, 256] (a fixed LCG), uniform entries, the
//            default size-class table, replayed from a HIP graph (n unused);
// c5         the batched mixed solve: 1024 matrices with orders in
//            [256, 2048], uniform entries with a zero diagonal on every 7th
//            index, 4 right-hand sides, tolerance 1e-12 (n unused);
// helpers    the kernels.h counterparts on large operands: ScaleOneByOne
//            over n^2 / 4 elements, ScaleTwoByTwo over n^2 / 8 pairs,
//            UpdateTrailingBlock n x n x 128.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bbk_ldlt.h"

namespace {

#define CHECK(x)                                                       \
  do {                                                                 \
    if ((x) != hipSuccess) {                                           \
      std::fprintf(stderr, "HIP call failed at line %d\n", __LINE__);  \
      return 1;                                                        \
    }                                                                  \
  } while (0)

std::vector<int> lcg_sizes(int count, int lo, int hi, uint64_t seed) {
  std::vector<int> ns(count);
  uint64_t x = seed;
  for (int i = 0; i < count; ++i) {
    x = x * 6364136223846793005ull + 1442695040888963407ull;
    ns[i] = lo + static_cast<int>((x >> 33) % static_cast<uint64_t>(hi - lo + 1));
  }
  return ns;
}

// Device pointer arrays of a batch carved from one allocation.
struct Batch {
  void* mem = nullptr;
  std::vector<int> ns;
  void** dA = nullptr;
  int64_t* dlda = nullptr;
  int* dn = nullptr;
  void** dP = nullptr;
  void** dV = nullptr;
};

int64_t padded_ld(int n) {
  int64_t ld = (n + 63) / 64 * 64;
  return ld % 1024 == 0 ? ld + 64 : ld;
}

// padded: leading dimensions padded_ld(n) (else n)
int make_batch(Batch* b, const std::vector<int>& ns, int elem, bool padded) {
  const int count = static_cast<int>(ns.size());
  b->ns = ns;
  std::vector<size_t> off(count), ioff(count);
  std::vector<int64_t> ld(count);
  size_t el = 0, il = 0;
  for (int i = 0; i < count; ++i) {
    ld[i] = padded ? padded_ld(ns[i]) : ns[i];
    off[i] = el;
    el += static_cast<size_t>(ld[i]) * ns[i];
    el = (el + 63) / 64 * 64;
    ioff[i] = il;
    il += 2 * static_cast<size_t>(ns[i]);
  }
  const size_t bytes = el * elem + il * sizeof(int) + count * (3 * sizeof(void*) + 12) + 4096;
  CHECK(hipMalloc(&b->mem, bytes));
  CHECK(hipMemset(b->mem, 0, bytes));
  char* base = static_cast<char*>(b->mem);
  int* ib = reinterpret_cast<int*>(base + el * elem);
  char* pb = reinterpret_cast<char*>(ib + il);
  pb = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(pb) + 255) / 256 * 256);
  b->dA = reinterpret_cast<void**>(pb);
  b->dP = b->dA + count;
  b->dV = b->dP + count;
  b->dlda = reinterpret_cast<int64_t*>(b->dV + count);
  b->dn = reinterpret_cast<int*>(b->dlda + count);
  std::vector<void*> hA(count), hP(count), hV(count);
  for (int i = 0; i < count; ++i) {
    hA[i] = base + off[i] * elem;
    hP[i] = ib + ioff[i];
    hV[i] = ib + ioff[i] + ns[i];
  }
  CHECK(hipMemcpy(b->dA, hA.data(), count * sizeof(void*), hipMemcpyHostToDevice));
  CHECK(hipMemcpy(b->dP, hP.data(), count * sizeof(void*), hipMemcpyHostToDevice));
  CHECK(hipMemcpy(b->dV, hV.data(), count * sizeof(void*), hipMemcpyHostToDevice));
  CHECK(hipMemcpy(b->dlda, ld.data(), count * sizeof(int64_t), hipMemcpyHostToDevice));
  CHECK(hipMemcpy(b->dn, ns.data(), count * sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

int run_c4(bbk_handle_t h, int reps) {
  const std::vector<int> ns = lcg_sizes(4096, 16, 256, 4096);
  for (int pass = 0; pass < 2; ++pass) {
    const bool f64 = pass == 0;
    Batch b;
    if (make_batch(&b, ns, f64 ? 8 : 4, false) != 0) return 1;
    const int count = static_cast<int>(ns.size());
    bbk_graph_t g = nullptr;
    int rc = f64 ? bbk_graph_create_dsytrf_batched(h, count, ns.data(),
                                                   reinterpret_cast<double* const*>(b.dA),
                                                   b.dlda, reinterpret_cast<int* const*>(b.dP),
                                                   reinterpret_cast<int* const*>(b.dV), &g)
                 : bbk_graph_create_ssytrf_batched(h, count, ns.data(),
                                                   reinterpret_cast<float* const*>(b.dA),
                                                   b.dlda, reinterpret_cast<int* const*>(b.dP),
                                                   reinterpret_cast<int* const*>(b.dV), &g);
    for (int r = 0; r < reps && rc == 0; ++r) {
      rc = f64 ? bbk_dgen_synth_batched(h, count, b.dn, 4096, 1.0, 0,
                                        reinterpret_cast<double* const*>(b.dA), b.dlda)
               : bbk_sgen_synth_batched(h, count, b.dn, 4096, 1.0, 0,
                                        reinterpret_cast<float* const*>(b.dA), b.dlda);
      if (rc == 0) rc = bbk_synchronize(h);
      if (rc == 0) rc = bbk_graph_launch(g);
      if (rc == 0) rc = bbk_synchronize(h);
      std::printf("c4 %s rep %d rc %d\n", f64 ? "f64" : "f32", r, rc);
      std::fflush(stdout);
    }
    if (g) bbk_graph_destroy(g);
    CHECK(hipFree(b.mem));
    if (rc != 0) return rc;
  }
  return 0;
}

int run_c5(bbk_handle_t h, int reps) {
  const std::vector<int> ns = lcg_sizes(1024, 256, 2048, 1024);
  const int count = static_cast<int>(ns.size());
  Batch a;
  if (make_batch(&a, ns, 8, true) != 0) return 1;
  // right-hand sides and solutions: n x 4 each
  size_t el = 0;
  std::vector<size_t> off(count);
  for (int i = 0; i < count; ++i) {
    off[i] = el;
    el += 4 * static_cast<size_t>(ns[i]);
  }
  double *B = nullptr, *X = nullptr;
  CHECK(hipMalloc(&B, el * 8));
  CHECK(hipMalloc(&X, el * 8));
  std::vector<double> hb(el);
  uint64_t x = 77;
  for (size_t i = 0; i < el; ++i) {
    x = x * 6364136223846793005ull + 1442695040888963407ull;
    hb[i] = static_cast<double>((x >> 11) & ((1ull << 52) - 1)) * 0x1.0p-52 * 2.0 - 1.0;
  }
  CHECK(hipMemcpy(B, hb.data(), el * 8, hipMemcpyHostToDevice));
  std::vector<void*> hB(count), hX(count);
  std::vector<int64_t> ld4(count, 4);
  for (int i = 0; i < count; ++i) {
    hB[i] = B + off[i];
    hX[i] = X + off[i];
  }
  void** dB = nullptr;
  void** dX = nullptr;
  int64_t* dld4 = nullptr;
  int* iters = nullptr;
  double* berr = nullptr;
  CHECK(hipMalloc(&dB, count * sizeof(void*)));
  CHECK(hipMalloc(&dX, count * sizeof(void*)));
  CHECK(hipMalloc(&dld4, count * sizeof(int64_t)));
  CHECK(hipMalloc(&iters, count * sizeof(int)));
  CHECK(hipMalloc(&berr, count * sizeof(double)));
  CHECK(hipMemcpy(dB, hB.data(), count * sizeof(void*), hipMemcpyHostToDevice));
  CHECK(hipMemcpy(dX, hX.data(), count * sizeof(void*), hipMemcpyHostToDevice));
  CHECK(hipMemcpy(dld4, ld4.data(), count * sizeof(int64_t), hipMemcpyHostToDevice));
  int rc = bbk_dgen_synth_batched(h, count, a.dn, 1024, 1.0, 7,
                                  reinterpret_cast<double* const*>(a.dA), a.dlda);
  if (rc == 0) rc = bbk_synchronize(h);
  for (int r = 0; r < reps && rc == 0; ++r) {
    rc = bbk_dsysv_mixed_batched(h, count, ns.data(), 4,
                                 reinterpret_cast<const double* const*>(a.dA), a.dlda,
                                 reinterpret_cast<const double* const*>(dB), dld4,
                                 reinterpret_cast<double* const*>(dX), dld4, 1e-12, 30, iters,
                                 berr);
    if (rc == BBK_ERR_NOT_CONVERGED) rc = 0;  // a profile of the kernels either way
    std::printf("c5 rep %d rc %d\n", r, rc);
    std::fflush(stdout);
  }
  CHECK(hipFree(dB));
  CHECK(hipFree(dX));
  CHECK(hipFree(dld4));
  CHECK(hipFree(iters));
  CHECK(hipFree(berr));
  CHECK(hipFree(B));
  CHECK(hipFree(X));
  CHECK(hipFree(a.mem));
  return rc;
}

int run_helpers(bbk_handle_t h, int n, int reps) {
  const size_t cnt = static_cast<size_t>(n) * n / 4;
  double *s1 = nullptr, *s2 = nullptr, *d1 = nullptr, *d2 = nullptr;
  CHECK(hipMalloc(&s1, cnt * 8));
  CHECK(hipMalloc(&s2, cnt * 8));
  CHECK(hipMalloc(&d1, cnt * 8));
  CHECK(hipMalloc(&d2, cnt * 8));
  const int inner = 128;
  double *left = nullptr, *work = nullptr, *target = nullptr;
  CHECK(hipMalloc(&left, static_cast<size_t>(inner) * n * 8));
  CHECK(hipMalloc(&work, static_cast<size_t>(inner) * n * 8));
  CHECK(hipMalloc(&target, static_cast<size_t>(n) * n * 8));
  CHECK(hipMemset(left, 0, static_cast<size_t>(inner) * n * 8));
  CHECK(hipMemset(work, 0, static_cast<size_t>(inner) * n * 8));
  // operands: the synthetic generator's values
  int rc = bbk_dgen_synth(h, n, 5, 1.0, 0, target, n);
  if (rc == 0) rc = bbk_dgen_synth(h, inner, 6, 1.0, 0, left, n);
  if (rc == 0) rc = bbk_dgen_synth(h, inner, 7, 1.0, 0, work, n);
  if (rc == 0) rc = bbk_synchronize(h);
  CHECK(hipMemcpy(s1, target, cnt * 8, hipMemcpyDeviceToDevice));
  CHECK(hipMemcpy(s2, target + cnt, cnt * 8, hipMemcpyDeviceToDevice));
  const int c1 = static_cast<int>(cnt), c2 = static_cast<int>(cnt / 2);
  for (int r = 0; r < reps && rc == 0; ++r) {
    rc = bbk_kernel_scale_1x1(h, c1, 1.75, s1, 1, d1, 1);
    if (rc == 0)
      rc = bbk_kernel_scale_2x2(h, c2, 2.0, 0.5, -3.0, s1, 1, s2, 1, d1, 1, d2, 1);
    if (rc == 0)
      rc = bbk_kernel_update_trailing(h, n, n, inner, left, n, work, n, target, n);
    if (rc == 0) rc = bbk_synchronize(h);
    std::printf("helpers rep %d rc %d\n", r, rc);
    std::fflush(stdout);
  }
  CHECK(hipFree(s1));
  CHECK(hipFree(s2));
  CHECK(hipFree(d1));
  CHECK(hipFree(d2));
  CHECK(hipFree(left));
  CHECK(hipFree(work));
  CHECK(hipFree(target));
  return rc;
}

}  // namespace

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 4096;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 1;
  const char* mode = argc > 3 ? argv[3] : "f64";
  const bool f32 = std::strcmp(mode, "f32") == 0;
  const bool mixed = std::strcmp(mode, "mixed") == 0;
  const bool sysv = std::strcmp(mode, "sysv") == 0;
  const int nb = argc > 4 ? std::atoi(argv[4]) : 0;
  bbk_handle_t h;
  if (bbk_create(&h, nullptr) != BBK_OK) return 1;
  int rc = 0;
  if (std::strcmp(mode, "c4") == 0 || std::strcmp(mode, "c5") == 0 ||
      std::strcmp(mode, "helpers") == 0) {
    rc = mode[0] == 'c' ? (mode[1] == '4' ? run_c4(h, reps) : run_c5(h, reps))
                        : run_helpers(h, n, reps);
  } else {
    int64_t lda = (n + 63) / 64 * 64;
    if (lda % 1024 == 0) lda += 64;
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
      CHECK(hipMemcpy(b, hb, 8 * (size_t)n, hipMemcpyHostToDevice));
      std::free(hb);
    }
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
        CHECK(hipMemcpy(hd, A0, bytes, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < (size_t)lda * n; ++i) tmp[i] = (float)hd[i];
        CHECK(hipMemcpy(S, tmp, bytes / 2, hipMemcpyHostToDevice));
        std::free(tmp);
        std::free(hd);
        rc = bbk_ssytrf(h, n, nb, S, lda, perm, piv);
      } else {
        CHECK(hipMemcpy(A, A0, bytes, hipMemcpyDeviceToDevice));
        rc = bbk_dsytrf(h, n, nb, A, lda, perm, piv);
      }
      if (rc == 0) rc = bbk_synchronize(h);
      std::printf("rep %d rc %d\n", r, rc);
      std::fflush(stdout);
    }
    CHECK(hipFree(A0));
    CHECK(hipFree(A));
    CHECK(hipFree(S));
    CHECK(hipFree(perm));
    CHECK(hipFree(piv));
    CHECK(hipFree(b));
    CHECK(hipFree(x));
  }
  if (rc != 0) std::fprintf(stderr, "error: %s\n", bbk_last_error(h));
  bbk_destroy(h);
  // every queue idle before the runtime and any attached tool tear down
  CHECK(hipDeviceSynchronize());
  std::printf("done\n");
  std::fflush(stdout);
  return rc;
}

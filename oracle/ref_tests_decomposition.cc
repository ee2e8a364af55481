// ref_tests_decomposition.cc -- C entry points over helpers of the
// reference's own test program tests/block_ldlt_decomposition_test.cc,
// compiled from that file where it lies (oracle/Makefile puts the
// reference's tests/ directory on the include path) into
// oracle/_ref/libbkref.so.  TEST INFRASTRUCTURE: it lets the tests pin the
// oracle's input generator against the reference's generator itself.
//
// The test program's helpers live in an anonymous namespace, so they are
// reachable from this translation unit only; its main() is renamed so the
// library has no entry point.
#define main bbk_ref_decomposition_test_main
#include "block_ldlt_decomposition_test.cc"
#undef main

#include <cstddef>

extern "C" {

// RandomSymmetricMatrix(size, seed, value_scale),
// tests/block_ldlt_decomposition_test.cc:148-161.  out: n*n row-major.
void ref_random_symmetric_matrix(int n, int seed, double scale, double* out) {
  const RegularMatrix m = RandomSymmetricMatrix(n, seed, scale);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) out[static_cast<std::size_t>(i) * n + j] = m(i, j);
}

}  // extern "C"

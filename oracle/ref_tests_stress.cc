// ref_tests_stress.cc -- C entry points over helpers of the reference's own
// large stress test tests/block_ldlt_large_stress_test.cc, compiled from
// that file where it lies (oracle/Makefile puts the reference's tests/
// directory on the include path) into oracle/_ref/libbkref.so.  TEST
// INFRASTRUCTURE: it pins the oracle's stress-matrix generator, residual
// sample positions and sampled residual against the reference's own code.
//
// The test program's helpers live in an anonymous namespace, so they are
// reachable from this translation unit only; its main() is renamed so the
// library has no entry point.
#define main bbk_ref_stress_test_main
#include "block_ldlt_large_stress_test.cc"
#undef main

#include <cstddef>
#include <cstring>

extern "C" {

// ResidualSamples(size, seed), tests/block_ldlt_large_stress_test.cc:204-218.
// rows/cols: room for kResidualSampleCount entries.  Returns the count.
int ref_residual_samples(int n, int seed, int* rows, int* cols) {
  const std::vector<std::pair<int, int>> s = ResidualSamples(n, seed);
  for (std::size_t t = 0; t < s.size(); ++t) {
    rows[t] = s[t].first;
    cols[t] = s[t].second;
  }
  return static_cast<int>(s.size());
}

// RandomSymmetricBlockMatrix(case), tests/block_ldlt_large_stress_test.cc:
// 116-136, read back through the BlockMatrix's symmetric accessor.
// out: n*n row-major.
void ref_random_symmetric_block_matrix(int n, int bs, int seed, double scale,
                                       double* out) {
  StressCase c;
  c.size = n;
  c.block_size = bs;
  c.value_scale = scale;
  c.seed = seed;
  const BlockMatrix m = RandomSymmetricBlockMatrix(c);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) out[static_cast<std::size_t>(i) * n + j] = m(i, j);
}

// The stress test's own case (RunCase, :246-265) without its printing:
// RandomSymmetricBlockMatrix, a kCompactOnly Factorize(), then
// SampledRelativeResidual(matrix, decomposition, seed) (:220-244).
// Outputs the compact factors (n*n upper, row-major), Permutation(),
// PivotTypes() and the sampled residual.  Returns 0, or 1 if it threw.
int ref_stress_case(int n, int bs, int seed, double scale, double* compact,
                    int* perm, int* piv, double* sampled_residual) {
  try {
    StressCase c;
    c.size = n;
    c.block_size = bs;
    c.value_scale = scale;
    c.seed = seed;
    const BlockMatrix matrix = RandomSymmetricBlockMatrix(c);
    BlockLdltDecomposition dec(
        matrix, BlockLdltDecomposition::FactorizationMode::kCompactOnly);
    dec.Factorize();
    *sampled_residual = SampledRelativeResidual(matrix, dec, seed);
    const naive_block_ldlt::RegularMatrix up = dec.CompactFactors().ToUpperRegular();
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        compact[static_cast<std::size_t>(i) * n + j] = up(i, j);
    for (int i = 0; i < n; ++i) {
      perm[i] = dec.Permutation()[i];
      piv[i] = dec.PivotTypes()[i];
    }
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

}  // extern "C"

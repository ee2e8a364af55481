// bk_gen.cc -- input generators for the oracle and the tests (TEST
// INFRASTRUCTURE).  The reference's test matrices come from libstdc++'s
// std::mt19937 + std::normal_distribution in a fixed draw order; calling
// the same library here regenerates them bit for bit on any host with the
// same libstdc++.  The synthetic workload generator (SplitMix64 keyed by
// (seed, i, j)) is an independent restatement of the formula the HIP
// library's generator kernel implements.
// Pins (tests/test_oracle.py): the two reference-matrix generators and the
// residual-sample positions are bitwise equal to the reference's own
// RandomSymmetricMatrix, RandomSymmetricBlockMatrix and ResidualSamples,
// compiled from its test sources (oracle/ref_tests_*.cc), and to fixtures
// those wrote; the SplitMix64 formula against an independent Python
// evaluation.
#include <cstddef>
#include <cstdint>
#include <random>

extern "C" {

// RandomSymmetricMatrix, tests/block_ldlt_decomposition_test.cc:148-161:
// upper triangle drawn row by row, mirrored.  out: n*n row-major.
void bko_gen_reference_symmetric(int n, int seed, double scale, double* out) {
  std::mt19937 generator(seed);
  std::normal_distribution<double> distribution(0.0, 1.0);
  for (int i = 0; i < n; ++i) {
    for (int j = i; j < n; ++j) {
      const double value = scale * distribution(generator);
      out[static_cast<std::size_t>(i) * n + j] = value;
      out[static_cast<std::size_t>(j) * n + i] = value;
    }
  }
}

// RandomSymmetricBlockMatrix, tests/block_ldlt_large_stress_test.cc:116-136:
// draws in (block_col, block_row <= block_col, row, col) order over the
// initial block geometry (src/block_matrix.cc:215-229), upper entries only.
void bko_gen_stress_matrix(int n, int bs, int seed, double scale,
                           double* out) {
  std::mt19937 generator(seed);
  std::normal_distribution<double> distribution(0.0, 1.0);
  const int nb = (n + bs - 1) / bs;
  for (int bc = 0; bc < nb; ++bc) {
    const int col0 = bc * bs;
    const int cols = (col0 + bs < n ? bs : n - col0);
    for (int br = 0; br <= bc; ++br) {
      const int row0 = br * bs;
      const int rows = (row0 + bs < n ? bs : n - row0);
      for (int r = 0; r < rows; ++r) {
        const int gr = row0 + r;
        for (int c = 0; c < cols; ++c) {
          const int gc = col0 + c;
          if (gr > gc) continue;
          const double value = scale * distribution(generator);
          out[static_cast<std::size_t>(gr) * n + gc] = value;
          out[static_cast<std::size_t>(gc) * n + gr] = value;
        }
      }
    }
  }
}

// ResidualSamples, tests/block_ldlt_large_stress_test.cc:204-218.
// rows/cols: count (4096 in the reference) entries each.
void bko_residual_samples(int n, int seed, int count, int* rows, int* cols) {
  int t = 0;
  const int fixed[4][2] = {{0, 0}, {0, n - 1}, {n / 2, n / 2}, {n - 1, n - 1}};
  for (; t < 4 && t < count; ++t) {
    rows[t] = fixed[t][0];
    cols[t] = fixed[t][1];
  }
  std::mt19937 generator(seed + 170003);
  std::uniform_int_distribution<int> distribution(0, n - 1);
  while (t < count) {
    // The reference builds the pair with a braced initializer, whose
    // elements are evaluated left to right: row first, then column.
    const int r = distribution(generator);
    const int c = distribution(generator);
    rows[t] = r;
    cols[t] = c;
    ++t;
  }
}

// Synthetic workload entries: u(seed, i, j) in [-1, 1), exact in binary64.
static inline std::uint64_t splitmix64(std::uint64_t x) {
  std::uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double bko_synth_uniform(std::uint64_t seed, int i, int j) {
  const std::uint64_t key =
      splitmix64(splitmix64(seed) ^ (static_cast<std::uint64_t>(i) << 32 |
                                     static_cast<std::uint32_t>(j)));
  return 2.0 * (static_cast<double>(key >> 11) * 0x1.0p-53) - 1.0;
}

// Symmetric synthetic matrix: A(i,j) = A(j,i) = scale * u(seed, min, max);
// A(i,i) = 0 when zero_diag_stride > 0 and i % zero_diag_stride == 0.
void bko_gen_synth_symmetric(int n, std::uint64_t seed, double scale,
                             int zero_diag_stride, double* out) {
  for (int i = 0; i < n; ++i) {
    for (int j = i; j < n; ++j) {
      double v = scale * bko_synth_uniform(seed, i, j);
      if (i == j && zero_diag_stride > 0 && i % zero_diag_stride == 0) v = 0.0;
      out[static_cast<std::size_t>(i) * n + j] = v;
      out[static_cast<std::size_t>(j) * n + i] = v;
    }
  }
}

}  // extern "C"

// ref_driver.cc -- C entry points over the reference library, compiled
// together with the reference's own sources (oracle/Makefile) into
// oracle/_ref/libbkref.so.  TEST INFRASTRUCTURE: it lets the tests pin the
// oracle against the real reference and lets the CPU baseline time the
// reference's Factorize().  It calls only the reference's public API
// (include/*.h under the reference tree).
#include <omp.h>

#include <chrono>
#include <cstddef>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "block.h"
#include "block_ldlt_decomposition.h"
#include "block_matrix.h"
#include "kernels.h"
#include "regular_matrix.h"

using naive_block_ldlt::Block;
using naive_block_ldlt::BlockLdltDecomposition;
using naive_block_ldlt::BlockMatrix;
using naive_block_ldlt::MatrixTriangle;
using naive_block_ldlt::RegularMatrix;

namespace {

void ExportGeometry(const BlockMatrix& m, int* geom) {
  // geom: for every block id (I + J(J+1)/2), {row0, col0, rows, cols, ld}.
  const int nb = m.num_blocks();
  for (int J = 0; J < nb; ++J)
    for (int I = 0; I <= J; ++I) {
      const Block& b = m.GetBlock(I, J);
      int* g = geom + 5 * m.BlockId(I, J);
      g[0] = b.row0();
      g[1] = b.col0();
      g[2] = b.rows();
      g[3] = b.cols();
      g[4] = b.leading_dimension();
    }
}

// The block-fill ingest of tests/block_ldlt_large_stress_test.cc:116-136:
// A (n*n row-major, upper triangle read) straight into the blocks, skipping
// the O(n^2 nb^2) element-wise constructor.
void FillBlocks(BlockMatrix& bm, const double* A, int n) {
  for (int J = 0; J < bm.num_blocks(); ++J)
    for (int I = 0; I <= J; ++I) {
      Block& b = bm.GetBlock(I, J);
      for (int r = 0; r < b.rows(); ++r)
        for (int c = 0; c < b.cols(); ++c) {
          const int gr = b.row0() + r, gc = b.col0() + c;
          if (gr > gc) continue;
          b(r, c) = A[static_cast<std::size_t>(gr) * n + gc];
        }
    }
}

double RelativeError(const RegularMatrix& actual,
                     const RegularMatrix& expected) {
  const double den = std::max(FrobeniusNorm(expected), 1.0);
  return FrobeniusNorm(Subtract(actual, expected)) / den;
}

}  // namespace

extern "C" {

// Full pipeline as the reference tests run it: RegularMatrix ->
// BlockMatrix(matrix, bs, triangle) -> BlockLdltDecomposition(mode) ->
// Factorize().  A: n*n row-major.  Outputs: compact upper (n*n, lower part
// zero), perm, piv, geometry (5 ints per block), workspace upper (n*n) and
// the full reconstruction residual (kFull only, else -1).  Returns 0, or 1
// if the reference threw.
int ref_factorize(int n, int bs, const double* A, int lower_input,
                  int full_mode, double* compact, int* perm, int* piv,
                  int* geom, double* workspace, double* residual) {
  try {
    RegularMatrix reg(n, n);
    std::memcpy(reg.Data(), A, sizeof(double) * static_cast<std::size_t>(n) * n);
    BlockMatrix bm(reg, bs,
                   lower_input ? MatrixTriangle::kLower : MatrixTriangle::kUpper);
    BlockLdltDecomposition dec(
        bm, full_mode ? BlockLdltDecomposition::FactorizationMode::kFull
                      : BlockLdltDecomposition::FactorizationMode::kCompactOnly);
    dec.Factorize();
    RegularMatrix up = dec.CompactFactors().ToUpperRegular();
    std::memcpy(compact, up.Data(), sizeof(double) * static_cast<std::size_t>(n) * n);
    for (int i = 0; i < n; ++i) {
      perm[i] = dec.Permutation()[i];
      piv[i] = dec.PivotTypes()[i];
    }
    if (geom) ExportGeometry(dec.CompactFactors(), geom);
    if (workspace) {
      RegularMatrix wup = dec.Workspace().ToUpperRegular();
      std::memcpy(workspace, wup.Data(),
                  sizeof(double) * static_cast<std::size_t>(n) * n);
    }
    if (residual) {
      if (full_mode) {
        RegularMatrix sym(n, n);
        // The symmetric matrix the chosen triangle describes.
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j) {
            const int lo = i < j ? i : j, hi = i < j ? j : i;
            sym(i, j) = lower_input ? A[static_cast<std::size_t>(hi) * n + lo]
                                    : A[static_cast<std::size_t>(lo) * n + hi];
          }
        *residual = RelativeError(dec.ReconstructOriginal(), sym);
      } else {
        *residual = -1.0;
      }
    }
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

// Compact-only factorization with the block-fill ingest of
// tests/block_ldlt_large_stress_test.cc:116-136 (skips the O(n^2 nb^2)
// element-wise constructor); times Factorize() alone with steady_clock.
// A: n*n row-major (upper triangle read).  compact may be NULL.
int ref_time_factorize(int n, int bs, const double* A, double* seconds,
                       double* compact, int* perm, int* piv) {
  try {
    BlockMatrix bm(n, bs);
    FillBlocks(bm, A, n);
    BlockLdltDecomposition dec(
        bm, BlockLdltDecomposition::FactorizationMode::kCompactOnly);
    const auto t0 = std::chrono::steady_clock::now();
    dec.Factorize();
    const auto t1 = std::chrono::steady_clock::now();
    *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (compact) {
      RegularMatrix up = dec.CompactFactors().ToUpperRegular();
      std::memcpy(compact, up.Data(),
                  sizeof(double) * static_cast<std::size_t>(n) * n);
    }
    if (perm)
      for (int i = 0; i < n; ++i) perm[i] = dec.Permutation()[i];
    if (piv)
      for (int i = 0; i < n; ++i) piv[i] = dec.PivotTypes()[i];
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

// The batched CPU baseline: every matrix's block storage is built first
// (block-fill ingest), then Factorize() (kCompactOnly) runs under an outer
// `omp parallel for schedule(dynamic)` over the matrices with nested
// parallel regions serialized, so each factorization runs on one thread;
// only that loop is timed.  Matrix i: ns[i] x ns[i] row-major at
// data + offs[i].  Returns the number of factorizations that threw.
int ref_time_batch(int count, const int* ns, const double* data, const long long* offs, int bs,
                   double* seconds) {
  std::vector<std::unique_ptr<BlockLdltDecomposition>> decs(count);
  for (int i = 0; i < count; ++i) {
    BlockMatrix bm(ns[i], bs);
    FillBlocks(bm, data + offs[i], ns[i]);
    decs[i] = std::make_unique<BlockLdltDecomposition>(
        bm, BlockLdltDecomposition::FactorizationMode::kCompactOnly);
  }
  const int levels = omp_get_max_active_levels();
  omp_set_max_active_levels(1);
  int failures = 0;
  const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic) reduction(+ : failures)
  for (int i = 0; i < count; ++i) {
    try {
      decs[i]->Factorize();
    } catch (const std::exception&) {
      failures += 1;
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  omp_set_max_active_levels(levels);
  *seconds = std::chrono::duration<double>(t1 - t0).count();
  return failures;
}

void ref_kernel_scale_1x1(int count, double pivot, const double* s, int ss,
                          double* d, int ds) {
  naive_block_ldlt::kernels::ScaleOneByOnePivot(count, pivot, s, ss, d, ds);
}

void ref_kernel_scale_2x2(int count, double a, double b, double c,
                          const double* s1, int s1s, const double* s2, int s2s,
                          double* d1, int d1s, double* d2, int d2s) {
  naive_block_ldlt::kernels::ScaleTwoByTwoPivot(count, a, b, c, s1, s1s, s2,
                                                s2s, d1, d1s, d2, d2s);
}

void ref_kernel_update_trailing(int rows, int cols, int inner,
                                const double* l, int ls, const double* w,
                                int ws, double* t, int ts) {
  naive_block_ldlt::kernels::UpdateTrailingBlock(rows, cols, inner, l, ls, w,
                                                 ws, t, ts);
}

}  // extern "C"

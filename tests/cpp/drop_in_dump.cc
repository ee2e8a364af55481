// drop_in_dump -- drives the C++ drop-in API (naive_block_ldlt) the way
// the reference's tests do and dumps what they inspect, for the Python
// parity tests to compare with the reference's own outputs.
//   drop_in_dump <in.bin> <out.bin>
// in.bin : int32 n, int32 bs, int32 lower, int32 full, then n*n doubles
//          (row-major; only the chosen triangle is meaningful)
// out.bin: int32 status, n*n doubles ToUpperRegular(CompactFactors()),
//          n int32 Permutation(), n int32 PivotTypes(),
//          5 int32 per block (row0, col0, rows, cols, ld) in BlockId order,
//          double residual ||ReconstructOriginal() - A||_F / max(||A||_F, 1)
//          (kFull) or -1, n*n doubles ToUpperRegular(Workspace()).
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <vector>

#include "block_ldlt_decomposition.h"
#include "block_matrix.h"
#include "regular_matrix.h"

using naive_block_ldlt::BlockLdltDecomposition;
using naive_block_ldlt::BlockMatrix;
using naive_block_ldlt::MatrixTriangle;
using naive_block_ldlt::RegularMatrix;

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  int hdr[4];
  if (std::fread(hdr, sizeof(int), 4, in) != 4) return 2;
  const int n = hdr[0], bs = hdr[1];
  const bool lower = hdr[2] != 0, full = hdr[3] != 0;
  std::vector<double> a(static_cast<size_t>(n) * n);
  if (std::fread(a.data(), sizeof(double), a.size(), in) != a.size()) return 2;
  std::fclose(in);
  FILE* out = std::fopen(argv[2], "wb");
  if (!out) return 2;
  int status = 0;
  try {
    RegularMatrix A(n, n, a);
    BlockMatrix bm(A, bs, lower ? MatrixTriangle::kLower : MatrixTriangle::kUpper);
    BlockLdltDecomposition dec(
        bm, full ? BlockLdltDecomposition::FactorizationMode::kFull
                 : BlockLdltDecomposition::FactorizationMode::kCompactOnly);
    dec.Factorize();
    std::fwrite(&status, sizeof(int), 1, out);
    RegularMatrix up = dec.CompactFactors().ToUpperRegular();
    std::fwrite(up.Data(), sizeof(double), static_cast<size_t>(n) * n, out);
    std::fwrite(dec.Permutation().data(), sizeof(int), n, out);
    std::fwrite(dec.PivotTypes().data(), sizeof(int), n, out);
    const BlockMatrix& f = dec.CompactFactors();
    for (int J = 0; J < f.num_blocks(); ++J)
      for (int I = 0; I <= J; ++I) {
        const auto& b = f.GetBlock(I, J);
        int g[5] = {b.row0(), b.col0(), b.rows(), b.cols(), b.leading_dimension()};
        std::fwrite(g, sizeof(int), 5, out);
      }
    double res = -1.0;
    if (full) {
      RegularMatrix sym = bm.ToRegular();
      res = naive_block_ldlt::FrobeniusNorm(
                naive_block_ldlt::Subtract(dec.ReconstructOriginal(), sym)) /
            std::max(naive_block_ldlt::FrobeniusNorm(sym), 1.0);
    }
    std::fwrite(&res, sizeof(double), 1, out);
    RegularMatrix wup = dec.Workspace().ToUpperRegular();
    std::fwrite(wup.Data(), sizeof(double), static_cast<size_t>(n) * n, out);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "drop_in_dump: %s\n", e.what());
    status = 1;
    std::fwrite(&status, sizeof(int), 1, out);
  }
  std::fclose(out);
  return status;
}

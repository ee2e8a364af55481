// block_ldlt_decomposition.cc -- the drop-in factorization entry point.
//
// Factorize() packs the upper triangle into one dense row-major buffer,
// hands it to the GPU through the C ABI (bbk_dsytrf_host), and unpacks the
// compact factors into block storage whose geometry is the one the
// reference builds (src/block_ldlt_decomposition.cc:50-120,589-631): a
// panel whose last column starts a 2x2 pivot gives that column to the next
// panel.  That geometry depends only on the pivot types and the block
// size, so it is replayed here from PivotTypes().
#include "block_ldlt_decomposition.h"

#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "bbk_ldlt.h"

namespace naive_block_ldlt {
namespace {

bbk_handle_t SharedHandle() {
  static std::once_flag once;
  static bbk_handle_t handle = nullptr;
  static int status = 0;
  std::call_once(once, [] { status = bbk_create(&handle, nullptr); });
  if (status != BBK_OK || handle == nullptr)
    throw std::runtime_error("blocked Bunch-Kaufman: no usable HIP device");
  return handle;
}

std::mutex& CallMutex() {
  static std::mutex m;
  return m;
}

void Check(int rc, bbk_handle_t h, const char* what) {
  if (rc == BBK_OK) return;
  const std::string msg = std::string(what) + ": " + bbk_last_error(h);
  if (rc < 0) throw std::invalid_argument(msg);
  if (rc == BBK_ERR_ALLOC) throw std::bad_alloc();
  throw std::runtime_error(msg);
}

// Copies the upper triangle between block views and a dense n x n buffer.
void Pack(const BlockMatrix& m, std::vector<double>* dense) {
  const int n = m.Rows();
  for (int J = 0; J < m.num_blocks(); ++J)
    for (int I = 0; I <= J; ++I) {
      const Block& b = m.GetBlock(I, J);
      for (int r = 0; r < b.rows(); ++r)
        for (int c = 0; c < b.cols(); ++c) {
          const int i = b.row0() + r, j = b.col0() + c;
          if (i <= j) (*dense)[static_cast<std::size_t>(i) * n + j] = b(r, c);
        }
    }
}

void Unpack(const std::vector<double>& dense, BlockMatrix* m) {
  const int n = m->Rows();
  for (int J = 0; J < m->num_blocks(); ++J)
    for (int I = 0; I <= J; ++I) {
      Block& b = m->GetBlock(I, J);
      for (int r = 0; r < b.rows(); ++r)
        for (int c = 0; c < b.cols(); ++c) {
          const int i = b.row0() + r, j = b.col0() + c;
          b(r, c) = i <= j ? dense[static_cast<std::size_t>(i) * n + j] : 0.0;
        }
    }
}

}  // namespace

BlockLdltDecomposition::BlockLdltDecomposition(const BlockMatrix& matrix, FactorizationMode mode)
    : original_(mode == FactorizationMode::kFull ? matrix.ToRegular() : RegularMatrix()),
      factors_(matrix),
      workspace_(matrix.Rows(), matrix.block_size()),
      l_factor_(mode == FactorizationMode::kFull ? RegularMatrix::Identity(matrix.Rows())
                                                 : RegularMatrix()),
      d_factor_(mode == FactorizationMode::kFull ? RegularMatrix(matrix.Rows(), matrix.Rows())
                                                 : RegularMatrix()),
      permutation_(matrix.Rows()),
      pivot_types_(matrix.Rows(), 0),
      mode_(mode) {
  for (int i = 0; i < matrix.Rows(); ++i) permutation_[i] = i;
}

// The panel-boundary moves of ShiftBoundaryColumnToNextBlock
// (src/block_ldlt_decomposition.cc:589-631), replayed on a matrix's views.
void BlockLdltDecomposition::ApplyPanelGeometry(BlockMatrix* m) const {
  const int n = m->Rows();
  const int nb = m->num_blocks();
  int start = 0;
  for (int P = 0; P < nb; ++P) {
    const Block& diag = m->GetBlock(P, P);
    start = diag.row0();
    const int end = diag.row0() + diag.rows();
    int k = start;
    while (k < end) {
      if (pivot_types_[k] == 2) {
        if (k == end - 1 && P + 1 < nb) {
          for (int J = P + 1; J < nb; ++J) m->ExpandBlock(P + 1, J);
          m->AddColumnLeft(P, P + 1);
          for (int J = P + 1; J < nb; ++J) m->ShrinkBlock(P, J);
          m->ShrinkBlock(P, P);
          break;
        }
        k += 2;
      } else {
        k += 1;
      }
    }
  }
  (void)n;
}

void BlockLdltDecomposition::Factorize() {
  const int n = factors_.Rows();
  std::vector<double> dense(static_cast<std::size_t>(n) * n, 0.0);
  Pack(factors_, &dense);
  {
    std::lock_guard<std::mutex> lock(CallMutex());
    bbk_handle_t h = SharedHandle();
    Check(bbk_dsytrf_host(h, n, 0, dense.data(), n, permutation_.data(), pivot_types_.data()),
          h, "bbk_dsytrf_host");
  }
  ApplyPanelGeometry(&factors_);
  ApplyPanelGeometry(&workspace_);
  Unpack(dense, &factors_);

  // Unscaled rows D L^T of every pivot.
  std::vector<double> w(static_cast<std::size_t>(n) * n, 0.0);
  auto U = [&](int i, int j) { return dense[static_cast<std::size_t>(i) * n + j]; };
  for (int p = 0; p < n;) {
    double* wp = &w[static_cast<std::size_t>(p) * n];
    if (pivot_types_[p] == 2) {
      const double a = U(p, p), b = U(p, p + 1), c = U(p + 1, p + 1);
      double* wq = &w[static_cast<std::size_t>(p + 1) * n];
      wp[p] = a;
      wp[p + 1] = b;
      wq[p + 1] = c;
      for (int j = p + 2; j < n; ++j) {
        wp[j] = a * U(p, j) + b * U(p + 1, j);
        wq[j] = b * U(p, j) + c * U(p + 1, j);
      }
      p += 2;
    } else {
      const double d = U(p, p);
      wp[p] = d;
      for (int j = p + 1; j < n; ++j) wp[j] = d * U(p, j);
      p += 1;
    }
  }
  Unpack(w, &workspace_);

  if (mode_ == FactorizationMode::kFull) {
    // ExtractRegularFactorsFromCompact, src/block_ldlt_decomposition.cc:746-763
    l_factor_.Fill(0.0);
    d_factor_.Fill(0.0);
    for (int i = 0; i < n; ++i) {
      l_factor_(i, i) = 1.0;
      d_factor_(i, i) = U(i, i);
      for (int j = i + 1; j < n; ++j) {
        if (j == i + 1 && pivot_types_[i] == 2) {
          d_factor_(i, j) = U(i, j);
          d_factor_(j, i) = U(i, j);
        } else {
          l_factor_(j, i) = U(i, j);
        }
      }
    }
  }
  factorized_ = true;
}

RegularMatrix BlockLdltDecomposition::ReconstructOriginal() const {
  if (mode_ != FactorizationMode::kFull)
    throw std::logic_error("ReconstructOriginal() requires full factor storage");
  if (!factorized_)
    throw std::logic_error("Factorize() must be called before ReconstructOriginal()");
  // P L D L^T P^T formed on the GPU from the compact factors (the reference
  // multiplies the dense L, D, L^T on the host)
  const RegularMatrix compact = factors_.ToUpperRegular();
  RegularMatrix out(original_.Rows(), original_.Cols());
  std::lock_guard<std::mutex> lock(CallMutex());
  bbk_handle_t h = SharedHandle();
  Check(bbk_dreconstruct_host(h, out.Rows(), compact.Data(), compact.LeadingDimension(),
                              permutation_.data(), pivot_types_.data(), out.Data(),
                              out.LeadingDimension()),
        h, "bbk_dreconstruct_host");
  return out;
}

}  // namespace naive_block_ldlt

// block_ldlt_decomposition.h -- drop-in for the reference's factorization
// entry point (include/block_ldlt_decomposition.h:11-30).
//
// Factorize() runs the blocked Bunch-Kaufman factorization on the GPU
// (bbk_dsytrf_host) and returns the reference's outputs: compact factors
// in upper block storage with the block geometry the reference's
// 2x2-boundary shifts produce, Permutation() and PivotTypes(), and in kFull
// mode the dense L and D.  Workspace() holds the unscaled rows D L^T of
// every pivot (the reference's candidate-scratch entries are not kept).
#ifndef NAIVE_BLOCK_LDLT_BLOCK_LDLT_DECOMPOSITION_H_
#define NAIVE_BLOCK_LDLT_BLOCK_LDLT_DECOMPOSITION_H_

#include <vector>

#include "block_matrix.h"
#include "regular_matrix.h"

namespace naive_block_ldlt {

class BlockLdltDecomposition {
 public:
  enum class FactorizationMode {
    kFull,
    kCompactOnly,
  };

  explicit BlockLdltDecomposition(const BlockMatrix& matrix,
                                  FactorizationMode mode = FactorizationMode::kFull);

  void Factorize();
  RegularMatrix ReconstructOriginal() const;

  const RegularMatrix& LFactor() const { return l_factor_; }
  const RegularMatrix& DFactor() const { return d_factor_; }
  const BlockMatrix& CompactFactors() const { return factors_; }
  const BlockMatrix& Workspace() const { return workspace_; }
  const std::vector<int>& Permutation() const { return permutation_; }
  const std::vector<int>& PivotTypes() const { return pivot_types_; }

 private:
  void ApplyPanelGeometry(BlockMatrix* matrix) const;

  RegularMatrix original_;
  BlockMatrix factors_;
  BlockMatrix workspace_;
  RegularMatrix l_factor_;
  RegularMatrix d_factor_;
  std::vector<int> permutation_;
  std::vector<int> pivot_types_;
  FactorizationMode mode_ = FactorizationMode::kFull;
  bool factorized_ = false;
};

}  // namespace naive_block_ldlt

#endif  // NAIVE_BLOCK_LDLT_BLOCK_LDLT_DECOMPOSITION_H_

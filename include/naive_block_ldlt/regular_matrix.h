// regular_matrix.h -- drop-in for the reference's dense row-major matrix
// (include/regular_matrix.h:11-45): ld = cols, 64-byte aligned, zeroed;
// the free functions are host utilities used to check factorizations.
#ifndef NAIVE_BLOCK_LDLT_REGULAR_MATRIX_H_
#define NAIVE_BLOCK_LDLT_REGULAR_MATRIX_H_

#include <vector>

#include "aligned_buffer.h"
#include "matrix.h"

namespace naive_block_ldlt {

class RegularMatrix : public Matrix {
 public:
  RegularMatrix() = default;
  RegularMatrix(int rows, int cols);
  RegularMatrix(int rows, int cols, const std::vector<double>& values);
  RegularMatrix(const RegularMatrix& other);
  RegularMatrix& operator=(const RegularMatrix& other);
  RegularMatrix(RegularMatrix&& other) noexcept = default;
  RegularMatrix& operator=(RegularMatrix&& other) noexcept = default;

  int Rows() const override { return rows_; }
  int Cols() const override { return cols_; }
  int LeadingDimension() const { return ld_; }
  double* Data() { return storage_.data(); }
  const double* Data() const { return storage_.data(); }
  double& operator()(int row, int col) override { return storage_.data()[row * ld_ + col]; }
  const double& operator()(int row, int col) const override {
    return storage_.data()[row * ld_ + col];
  }
  void Fill(double value) { storage_.Fill(value); }
  static RegularMatrix Identity(int size);

 private:
  int rows_ = 0;
  int cols_ = 0;
  int ld_ = 0;
  AlignedBuffer storage_;
};

double FrobeniusNorm(const RegularMatrix& matrix);
RegularMatrix Multiply(const RegularMatrix& left, const RegularMatrix& right);
RegularMatrix Transpose(const RegularMatrix& matrix);
RegularMatrix Subtract(const RegularMatrix& left, const RegularMatrix& right);

}  // namespace naive_block_ldlt

#endif  // NAIVE_BLOCK_LDLT_REGULAR_MATRIX_H_

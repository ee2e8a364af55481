// regular_matrix.cc -- dense row-major matrix of the drop-in API and the
// host utilities used to check factorizations (FrobeniusNorm, Multiply,
// Transpose, Subtract).  Same contracts as the reference
// (src/regular_matrix.cc:14-148): shape errors throw
// std::invalid_argument, storage is zero-initialised, and the utilities
// sum in row-major i-k-j order, so they give the reference's values.
#include "regular_matrix.h"

#include <cmath>
#include <cstddef>
#include <stdexcept>

namespace naive_block_ldlt {
namespace {

std::size_t ElementCount(int rows, int cols) {
  if (rows < 0 || cols < 0) throw std::invalid_argument("matrix shape must be nonnegative");
  return static_cast<std::size_t>(rows) * static_cast<std::size_t>(cols);
}

}  // namespace

RegularMatrix::RegularMatrix(int rows, int cols)
    : rows_(rows), cols_(cols), ld_(cols), storage_(ElementCount(rows, cols)) {}

RegularMatrix::RegularMatrix(int rows, int cols, const std::vector<double>& values)
    : RegularMatrix(rows, cols) {
  if (values.size() != ElementCount(rows, cols))
    throw std::invalid_argument("value count does not match matrix shape");
  for (std::size_t i = 0; i < values.size(); ++i) storage_.data()[i] = values[i];
}

RegularMatrix::RegularMatrix(const RegularMatrix& other) : RegularMatrix(other.rows_, other.cols_) {
  for (std::size_t i = 0; i < storage_.size(); ++i) storage_.data()[i] = other.storage_.data()[i];
}

RegularMatrix& RegularMatrix::operator=(const RegularMatrix& other) {
  if (this == &other) return *this;
  rows_ = other.rows_;
  cols_ = other.cols_;
  ld_ = other.ld_;
  storage_.Resize(other.storage_.size());
  for (std::size_t i = 0; i < storage_.size(); ++i) storage_.data()[i] = other.storage_.data()[i];
  return *this;
}

RegularMatrix RegularMatrix::Identity(int size) {
  RegularMatrix id(size, size);
  for (int i = 0; i < size; ++i) id(i, i) = 1.0;
  return id;
}

double FrobeniusNorm(const RegularMatrix& matrix) {
  if (matrix.Rows() == 0 || matrix.Cols() == 0) return 0.0;
  double sum = 0.0;
  for (int i = 0; i < matrix.Rows(); ++i)
    for (int j = 0; j < matrix.Cols(); ++j) sum += matrix(i, j) * matrix(i, j);
  return std::sqrt(sum);
}

RegularMatrix Multiply(const RegularMatrix& left, const RegularMatrix& right) {
  if (left.Cols() != right.Rows())
    throw std::invalid_argument("incompatible matrix multiplication shapes");
  RegularMatrix out(left.Rows(), right.Cols());
  for (int i = 0; i < left.Rows(); ++i)
    for (int k = 0; k < left.Cols(); ++k) {
      const double a = left(i, k);
      for (int j = 0; j < right.Cols(); ++j) out(i, j) += a * right(k, j);
    }
  return out;
}

RegularMatrix Transpose(const RegularMatrix& matrix) {
  RegularMatrix out(matrix.Cols(), matrix.Rows());
  for (int i = 0; i < matrix.Rows(); ++i)
    for (int j = 0; j < matrix.Cols(); ++j) out(j, i) = matrix(i, j);
  return out;
}

RegularMatrix Subtract(const RegularMatrix& left, const RegularMatrix& right) {
  if (left.Rows() != right.Rows() || left.Cols() != right.Cols())
    throw std::invalid_argument("incompatible matrix subtraction shapes");
  RegularMatrix out(left.Rows(), left.Cols());
  for (int i = 0; i < left.Rows(); ++i)
    for (int j = 0; j < left.Cols(); ++j) out(i, j) = left(i, j) - right(i, j);
  return out;
}

}  // namespace naive_block_ldlt

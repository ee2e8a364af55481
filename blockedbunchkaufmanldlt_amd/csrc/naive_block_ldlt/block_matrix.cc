// block_matrix.cc -- upper-triangular block storage of the drop-in API.
// Geometry and error contracts follow the reference
// (src/block_matrix.cc:11-312): blocks (I <= J) of bs x bs, ragged last
// row/column, one spare row (I > 0) and column (J > 0) of capacity for the
// 2x2 boundary shift; resizes keep every element at its global position
// and zero the cells they expose.  Element lookup is O(1) from the block
// grid (checked against the live views) instead of a scan of all blocks.
#include "block_matrix.h"

#include <algorithm>
#include <cstddef>
#include <stdexcept>
#include <utility>

namespace naive_block_ldlt {

BlockMatrix::BlockMatrix(int size, int block_size)
    : size_(size),
      block_size_(block_size),
      num_blocks_(block_size > 0 ? (size + block_size - 1) / block_size : 0) {
  if (size <= 0) throw std::invalid_argument("matrix size must be positive");
  if (block_size <= 0) throw std::invalid_argument("block size must be positive");
  Allocate();
}

BlockMatrix::BlockMatrix(const RegularMatrix& matrix, int block_size,
                         MatrixTriangle input_triangle)
    : BlockMatrix(matrix.Rows(), block_size) {
  if (matrix.Rows() != matrix.Cols())
    throw std::invalid_argument("block matrix input must be square");
  const bool upper = input_triangle == MatrixTriangle::kUpper;
  for (int J = 0; J < num_blocks_; ++J)
    for (int I = 0; I <= J; ++I) {
      Block& b = blocks_[BlockId(I, J)];
      for (int r = 0; r < b.rows(); ++r)
        for (int c = 0; c < b.cols(); ++c) {
          const int i = b.row0() + r, j = b.col0() + c;
          if (i > j) continue;
          b(r, c) = upper ? matrix(i, j) : matrix(j, i);
        }
    }
}

BlockMatrix::BlockMatrix(const BlockMatrix& other)
    : size_(other.size_),
      block_size_(other.block_size_),
      num_blocks_(other.num_blocks_),
      row_capacity_(other.row_capacity_),
      col_capacity_(other.col_capacity_) {
  storage_.reserve(other.storage_.size());
  blocks_.reserve(other.blocks_.size());
  for (std::size_t id = 0; id < other.blocks_.size(); ++id) {
    const AlignedBuffer& src = other.storage_[id];
    storage_.emplace_back(src.size());
    std::copy(src.data(), src.data() + src.size(), storage_.back().data());
    const Block& b = other.blocks_[id];
    blocks_.emplace_back(storage_.back().data(), b.rows(), b.cols(), b.leading_dimension(),
                         b.row0(), b.col0(), b.block_row(), b.block_col());
  }
}

BlockMatrix& BlockMatrix::operator=(const BlockMatrix& other) {
  if (this != &other) {
    BlockMatrix tmp(other);
    *this = std::move(tmp);
  }
  return *this;
}

void BlockMatrix::Allocate() {
  const int count = num_blocks_ * (num_blocks_ + 1) / 2;
  storage_.reserve(count);
  blocks_.reserve(count);
  row_capacity_.reserve(count);
  col_capacity_.reserve(count);
  for (int J = 0; J < num_blocks_; ++J) {
    const int col0 = J * block_size_;
    const int cols = std::min(block_size_, size_ - col0);
    const int ccap = cols + (J > 0 ? 1 : 0);
    for (int I = 0; I <= J; ++I) {
      const int row0 = I * block_size_;
      const int rows = std::min(block_size_, size_ - row0);
      const int rcap = rows + (I > 0 ? 1 : 0);
      row_capacity_.push_back(rcap);
      col_capacity_.push_back(ccap);
      storage_.emplace_back(static_cast<std::size_t>(rcap) * static_cast<std::size_t>(ccap));
      blocks_.emplace_back(storage_.back().data(), rows, cols, ccap, row0, col0, I, J);
    }
  }
}

int BlockMatrix::BlockId(int block_row, int block_col) const {
  if (block_row > block_col) std::swap(block_row, block_col);
  return block_row + block_col * (block_col + 1) / 2;
}

Block& BlockMatrix::GetBlock(int block_row, int block_col) {
  return blocks_.at(BlockId(block_row, block_col));
}

const Block& BlockMatrix::GetBlock(int block_row, int block_col) const {
  return blocks_.at(BlockId(block_row, block_col));
}

static bool Contains(const Block& b, int row, int col) {
  return b.row0() <= row && row < b.row0() + b.rows() && b.col0() <= col &&
         col < b.col0() + b.cols();
}

const Block* BlockMatrix::Locate(int row, int col) const {
  if (row < 0 || col < 0 || row >= size_ || col >= size_ || num_blocks_ == 0) return nullptr;
  // A 2x2 boundary shift moves a view edge by at most one element, so the
  // element is in its home block or a neighbour of it.
  const int hi = row / block_size_, hj = col / block_size_;
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj) {
      const int I = hi + di, J = hj + dj;
      if (I < 0 || J < 0 || I >= num_blocks_ || J >= num_blocks_ || I > J) continue;
      const Block& b = blocks_[BlockId(I, J)];
      if (Contains(b, row, col)) return &b;
    }
  for (const Block& b : blocks_)
    if (Contains(b, row, col)) return &b;
  return nullptr;
}

double& BlockMatrix::operator()(int row, int col) {
  if (row > col) std::swap(row, col);
  const Block* b = Locate(row, col);
  if (b == nullptr) throw std::out_of_range("matrix index is outside active block views");
  Block& mb = blocks_[BlockId(b->block_row(), b->block_col())];
  return mb(row - mb.row0(), col - mb.col0());
}

const double& BlockMatrix::operator()(int row, int col) const {
  if (row > col) std::swap(row, col);
  const Block* b = Locate(row, col);
  if (b == nullptr) throw std::out_of_range("matrix index is outside active block views");
  return (*b)(row - b->row0(), col - b->col0());
}

BlockColumnLocation BlockMatrix::GlobalColumnToBlockColumn(int block_row, int preferred_block_col,
                                                           int column) const {
  auto holds = [&](int J) {
    const Block& b = GetBlock(block_row, J);
    return b.col0() <= column && column < b.col0() + b.cols();
  };
  if (holds(preferred_block_col))
    return {block_row, preferred_block_col, column - GetBlock(block_row, preferred_block_col).col0()};
  for (int J = preferred_block_col - 1; J >= 0; --J)
    if (holds(J)) return {block_row, J, column - GetBlock(block_row, J).col0()};
  for (int J = preferred_block_col + 1; J < num_blocks_; ++J)
    if (holds(J)) return {block_row, J, column - GetBlock(block_row, J).col0()};
  throw std::out_of_range("global column is outside active block row");
}

RegularMatrix BlockMatrix::ToRegular() const {
  RegularMatrix out(size_, size_);
  for (const Block& b : blocks_)
    for (int r = 0; r < b.rows(); ++r)
      for (int c = 0; c < b.cols(); ++c) {
        const int i = b.row0() + r, j = b.col0() + c;
        if (i < 0 || j < 0 || i >= size_ || j >= size_ || i > j) continue;
        out(i, j) = b(r, c);
        out(j, i) = b(r, c);
      }
  return out;
}

RegularMatrix BlockMatrix::ToUpperRegular() const {
  RegularMatrix out(size_, size_);
  for (const Block& b : blocks_)
    for (int r = 0; r < b.rows(); ++r)
      for (int c = 0; c < b.cols(); ++c) {
        const int i = b.row0() + r, j = b.col0() + c;
        if (i < 0 || j < 0 || i >= size_ || j >= size_ || i > j) continue;
        out(i, j) = b(r, c);
      }
  return out;
}

void BlockMatrix::ExpandBlock(int block_row, int block_col, int col_size) {
  const Block b = GetBlock(block_row, block_col);
  if (block_row == block_col)
    Reshape(block_row, block_col, b.rows() + 1, b.cols() + col_size, 1, col_size, b.row0() - 1,
            b.col0() - col_size);
  else
    Reshape(block_row, block_col, b.rows() + 1, b.cols(), 1, 0, b.row0() - 1, b.col0());
}

void BlockMatrix::ShrinkBlock(int block_row, int block_col) {
  const Block b = GetBlock(block_row, block_col);
  if (block_row == block_col) {
    if (b.rows() == 0 || b.cols() == 0) return;
    Reshape(block_row, block_col, b.rows() - 1, b.cols() - 1, 0, 0, b.row0(), b.col0());
  } else {
    if (b.rows() == 0) return;
    Reshape(block_row, block_col, b.rows() - 1, b.cols(), 0, 0, b.row0(), b.col0());
  }
}

void BlockMatrix::AddColumnLeft(int block_row, int block_col, int col_size) {
  const Block b = GetBlock(block_row, block_col);
  Reshape(block_row, block_col, b.rows(), b.cols() + col_size, 0, col_size, b.row0(),
          b.col0() - col_size);
}

// New view of `rows` x `cols` whose element (r, c) is the old element
// (r - row_shift, c - col_shift), zero where that is outside the old view.
void BlockMatrix::Reshape(int block_row, int block_col, int rows, int cols, int row_shift,
                          int col_shift, int row0, int col0) {
  const int id = BlockId(block_row, block_col);
  if (rows < 0 || cols < 0) throw std::invalid_argument("block dimensions must be non-negative");
  if (rows > row_capacity_.at(id) || cols > col_capacity_.at(id))
    throw std::logic_error("block resize exceeds preallocated capacity");
  Block& b = blocks_.at(id);
  const int ld = col_capacity_[id];
  std::vector<double> old(static_cast<std::size_t>(b.rows()) * b.cols());
  for (int r = 0; r < b.rows(); ++r)
    for (int c = 0; c < b.cols(); ++c) old[static_cast<std::size_t>(r) * b.cols() + c] = b(r, c);
  const int orows = b.rows(), ocols = b.cols();
  double* data = storage_[id].data();
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) {
      const int orr = r - row_shift, oc = c - col_shift;
      data[r * ld + c] = (orr >= 0 && orr < orows && oc >= 0 && oc < ocols)
                             ? old[static_cast<std::size_t>(orr) * ocols + oc]
                             : 0.0;
    }
  b.Reset(data, rows, cols, ld, row0, col0, block_row, block_col);
}

}  // namespace naive_block_ldlt

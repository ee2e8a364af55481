// aligned_buffer.h -- drop-in for the reference's include/aligned_buffer.h
// (AlignedBuffer, kDefaultAlignment; include/aligned_buffer.h:16-112):
// an owning, move-only, zero-initialised, 64-byte aligned array of double.
#ifndef NAIVE_BLOCK_LDLT_ALIGNED_BUFFER_H_
#define NAIVE_BLOCK_LDLT_ALIGNED_BUFFER_H_

#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <new>

namespace naive_block_ldlt {

inline constexpr std::size_t kDefaultAlignment = 64;

class AlignedBuffer {
 public:
  AlignedBuffer() = default;
  explicit AlignedBuffer(std::size_t size) : AlignedBuffer(size, kDefaultAlignment) {}
  AlignedBuffer(std::size_t size, std::size_t alignment) : alignment_(alignment) {
    Acquire(size);
  }
  AlignedBuffer(const AlignedBuffer&) = delete;
  AlignedBuffer& operator=(const AlignedBuffer&) = delete;
  AlignedBuffer(AlignedBuffer&& other) noexcept { Take(other); }
  AlignedBuffer& operator=(AlignedBuffer&& other) noexcept {
    if (this != &other) {
      Release();
      Take(other);
    }
    return *this;
  }
  ~AlignedBuffer() { Release(); }

  // Reallocates (contents zeroed) unless the size is unchanged.
  void Resize(std::size_t size) {
    if (size == size_) return;
    Release();
    Acquire(size);
  }
  void Fill(double value) {
    for (std::size_t i = 0; i < size_; ++i) data_[i] = value;
  }
  double* data() { return data_; }
  const double* data() const { return data_; }
  std::size_t size() const { return size_; }
  std::size_t alignment() const { return alignment_; }

 private:
  void Acquire(std::size_t size) {
    size_ = size;
    data_ = nullptr;
    if (size == 0) return;
    std::size_t bytes = size * sizeof(double);
    bytes = (bytes + alignment_ - 1) / alignment_ * alignment_;
    data_ = static_cast<double*>(std::aligned_alloc(alignment_, bytes));
    if (data_ == nullptr) throw std::bad_alloc();
    std::memset(data_, 0, bytes);
  }
  void Release() {
    std::free(data_);
    data_ = nullptr;
    size_ = 0;
  }
  void Take(AlignedBuffer& other) {
    data_ = other.data_;
    size_ = other.size_;
    alignment_ = other.alignment_;
    other.data_ = nullptr;
    other.size_ = 0;
  }

  double* data_ = nullptr;
  std::size_t size_ = 0;
  std::size_t alignment_ = kDefaultAlignment;
};

}  // namespace naive_block_ldlt

#endif  // NAIVE_BLOCK_LDLT_ALIGNED_BUFFER_H_

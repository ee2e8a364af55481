// bbk_handle.hpp -- state behind a bbk_handle_t.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <vector>

namespace bbk {

struct Handle {
  enum { kPanel = 0, kUpdate = 1, kSwap = 2, kGroups = 3 };

  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  void* ws = nullptr;       // device scratch, grown on demand
  size_t ws_bytes = 0;
  std::string last_error;
  bool timing = false;
  int panels = 0;
  std::vector<hipEvent_t> marks[kGroups];

  ~Handle();
  int fail(int code, const char* what);
  int hip_fail(hipError_t e, const char* what);
  int reserve(size_t bytes);
  void timing_reset();
  void timing_mark(int group, bool begin);
  double timing_total(int group);
};

int default_panel_width(int n, int elem_bytes);

template <typename T>
int factorize(Handle* h, int n, int nb, T* A, int64_t lda, int* perm,
              int* piv);

}  // namespace bbk

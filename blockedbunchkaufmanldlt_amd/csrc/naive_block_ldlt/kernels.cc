// kernels.cc -- the drop-in kernels:: helpers (include/kernels.h) on the
// GPU: operands are copied to the device, the gfx950 kernel runs, results
// come back (bbk_kernel_*_host).  Same argument meaning as
// src/kernels.cc:11-91; errors throw std::invalid_argument /
// std::runtime_error like the rest of the drop-in API.
#include "kernels.h"

#include <mutex>
#include <stdexcept>
#include <string>

#include "bbk_ldlt.h"

namespace naive_block_ldlt::kernels {
namespace {

bbk_handle_t Handle() {
  static std::once_flag once;
  static bbk_handle_t handle = nullptr;
  static int status = 0;
  std::call_once(once, [] { status = bbk_create(&handle, nullptr); });
  if (status != BBK_OK || handle == nullptr)
    throw std::runtime_error("blocked Bunch-Kaufman kernels: no usable HIP device");
  return handle;
}

std::mutex& Mutex() {
  static std::mutex m;
  return m;
}

void Check(int rc, bbk_handle_t h, const char* what) {
  if (rc == BBK_OK) return;
  const std::string msg = std::string(what) + ": " + bbk_last_error(h);
  if (rc < 0) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

}  // namespace

void ScaleOneByOnePivot(int count, double pivot, const double* source, int source_stride,
                        double* destination, int destination_stride) {
  if (count <= 0) return;
  std::lock_guard<std::mutex> lock(Mutex());
  bbk_handle_t h = Handle();
  Check(bbk_kernel_scale_1x1_host(h, count, pivot, source, source_stride, destination,
                                  destination_stride),
        h, "ScaleOneByOnePivot");
}

void ScaleTwoByTwoPivot(int count, double pivot_top_left, double pivot_off_diagonal,
                        double pivot_bottom_right, const double* first_source,
                        int first_source_stride, const double* second_source,
                        int second_source_stride, double* first_destination,
                        int first_destination_stride, double* second_destination,
                        int second_destination_stride) {
  if (count <= 0) return;
  std::lock_guard<std::mutex> lock(Mutex());
  bbk_handle_t h = Handle();
  Check(bbk_kernel_scale_2x2_host(h, count, pivot_top_left, pivot_off_diagonal,
                                  pivot_bottom_right, first_source, first_source_stride,
                                  second_source, second_source_stride, first_destination,
                                  first_destination_stride, second_destination,
                                  second_destination_stride),
        h, "ScaleTwoByTwoPivot");
}

void UpdateTrailingBlock(int rows, int cols, int inner_dim, const double* left_panel,
                         int left_panel_stride, const double* workspace_panel,
                         int workspace_panel_stride, double* target, int target_stride) {
  if (rows <= 0 || cols <= 0 || inner_dim <= 0) return;
  std::lock_guard<std::mutex> lock(Mutex());
  bbk_handle_t h = Handle();
  Check(bbk_kernel_update_trailing_host(h, rows, cols, inner_dim, left_panel, left_panel_stride,
                                        workspace_panel, workspace_panel_stride, target,
                                        target_stride),
        h, "UpdateTrailingBlock");
}

}  // namespace naive_block_ldlt::kernels

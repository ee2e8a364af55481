// bbk_solve.hip -- solves, refinement and batched entry points (in progress).
#include <hip/hip_runtime.h>
#include "../../include/bbk_ldlt.h"
#include "bbk_handle.hpp"
extern "C" {
int bbk_dsytrs(bbk_handle_t, int, int, const double*, int64_t, const int*, const int*, double*, int64_t) { return BBK_ERR_UNSUPPORTED; }
int bbk_ssytrs(bbk_handle_t, int, int, const float*, int64_t, const int*, const int*, float*, int64_t) { return BBK_ERR_UNSUPPORTED; }
int bbk_dsyberr(bbk_handle_t, int, int, const double*, int64_t, const double*, int64_t, const double*, int64_t, double*) { return BBK_ERR_UNSUPPORTED; }
int bbk_dsysv_mixed(bbk_handle_t, int, int, const double*, int64_t, const double*, int64_t, double*, int64_t, double, int, int*, double*) { return BBK_ERR_UNSUPPORTED; }
int bbk_dsysv(bbk_handle_t, int, int, const double*, int64_t, const double*, int64_t, double*, int64_t, double*) { return BBK_ERR_UNSUPPORTED; }
int bbk_dsytrf_batched(bbk_handle_t, int, const int*, double* const*, const int64_t*, int* const*, int* const*) { return BBK_ERR_UNSUPPORTED; }
int bbk_ssytrf_batched(bbk_handle_t, int, const int*, float* const*, const int64_t*, int* const*, int* const*) { return BBK_ERR_UNSUPPORTED; }
}

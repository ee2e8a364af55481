// bbk_lapack.hip -- LAPACK-convention interop for the compact factors:
// conversion between this library's (and the reference's) form and the
// form LAPACK dsytrf('L') returns, both stored in place in the same bytes
// (row-major upper triangle = column-major lower triangle).
//
// This library (src/block_ldlt_decomposition.cc:392-587, the interchanges
// are applied to the earlier L rows too):
//   A(perm[i], perm[j]) = (L D L^T)(i, j), L(i, k) = F(k, i) for i > k.
// LAPACK dsytrf('L'):
//   A = P(1) L(1) P(2) L(2) ... P(s) L(s) D (...)^T, where P(j) is the
//   interchange of step j (rows and columns a_j <-> b_j, a_j = j for a 1x1
//   pivot, j + 1 for a 2x2 one, b_j >= a_j) and L(j) holds step j's
//   multipliers as LAPACK stored them, before the later steps' interchanges.
// Moving the permutations left shows that step k's multipliers here are
// LAPACK's permuted by P(k+1) ... P(s); so the conversion applies the later
// steps' interchanges to the earlier columns of L, in reverse step order to
// go to LAPACK, in step order to come back.  The 2x2 pivot's D entry F(k,
// k+1) and the diagonal are untouched (later interchanges only move rows
// beyond k + 1).  The interchange sequence is unique: position a_j holds its
// final index after step j, so b_j is where that index stands at step j.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/bbk_ldlt.h"
#include "bbk_handle.hpp"

namespace bbk {

// Thread c applies, to F's row c (the multipliers L(., c)), the
// interchanges of the steps j > c: swap F(c, a_j) and F(c, b_j).  REVERSE:
// last step first (to LAPACK form); otherwise first step first (back).
// steps[3 t] = step j, [3 t + 1] = a_j, [3 t + 2] = b_j, sorted by j, only
// steps that interchange (a_j != b_j).
template <bool REVERSE>
__global__ void k_apply_later_swaps(int n, double* __restrict__ F, int64_t ldf,
                                    const int* __restrict__ steps, int nsteps) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double* row = F + (int64_t)c * ldf;
  // first listed step after c
  int lo = 0, hi = nsteps;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (steps[3 * mid] > c)
      hi = mid;
    else
      lo = mid + 1;
  }
  if (REVERSE) {
    for (int t = nsteps - 1; t >= lo; --t) {
      const int a = steps[3 * t + 1], b = steps[3 * t + 2];
      const double x = row[a];
      row[a] = row[b];
      row[b] = x;
    }
  } else {
    for (int t = lo; t < nsteps; ++t) {
      const int a = steps[3 * t + 1], b = steps[3 * t + 2];
      const double x = row[a];
      row[a] = row[b];
      row[b] = x;
    }
  }
}

static int upload_and_apply(Handle* h, int n, double* d_F, int64_t ldf,
                            const std::vector<int>& steps, bool reverse) {
  const int nsteps = static_cast<int>(steps.size() / 3);
  if (nsteps == 0) return BBK_OK;
  int* d_steps = nullptr;
  hipError_t e = hipMalloc(&d_steps, sizeof(int) * steps.size());
  if (e != hipSuccess) return h->hip_fail(e, "interchange list");
  e = hipMemcpyAsync(d_steps, steps.data(), sizeof(int) * steps.size(), hipMemcpyHostToDevice,
                     h->stream);
  if (e == hipSuccess) {
    const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
    if (reverse)
      hipLaunchKernelGGL(k_apply_later_swaps<true>, dim3(blocks), dim3(256), 0, h->stream, n, d_F,
                         ldf, d_steps, nsteps);
    else
      hipLaunchKernelGGL(k_apply_later_swaps<false>, dim3(blocks), dim3(256), 0, h->stream, n,
                         d_F, ldf, d_steps, nsteps);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(d_steps);
  if (e != hipSuccess) return h->hip_fail(e, "interchange kernel");
  return BBK_OK;
}

}  // namespace bbk

using bbk::Handle;

extern "C" {

int bbk_dsytrf_to_lapack(bbk_handle_t handle, int n, double* d_F, int64_t ldf,
                         const int* d_perm, const int* d_piv, int* d_ipiv) {
  if (!handle) return -1;
  if (n <= 0) return -2;
  if (!d_F) return -3;
  if (ldf < n) return -4;
  if (!d_perm) return -5;
  if (!d_piv) return -6;
  if (!d_ipiv) return -7;
  Handle* h = reinterpret_cast<Handle*>(handle);
  std::vector<int> perm(n), piv(n), ipiv(n);
  hipError_t e = hipMemcpyAsync(perm.data(), d_perm, sizeof(int) * n, hipMemcpyDeviceToHost,
                                h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(piv.data(), d_piv, sizeof(int) * n, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return h->hip_fail(e, "read pivots");
  // the interchange of every step, from the final permutation
  std::vector<int> cur(n), pos(n);
  for (int i = 0; i < n; ++i) cur[i] = pos[i] = i;
  std::vector<int> steps;
  for (int k = 0; k < n;) {
    const bool two = piv[k] == 2;
    if (two && (k + 1 >= n || piv[k + 1] != -2)) return h->fail(-6, "invalid 2x2 pivot encoding");
    if (!two && piv[k] != 1) return h->fail(-6, "invalid pivot type");
    const int a = two ? k + 1 : k;
    if (two && cur[k] != perm[k]) return h->fail(-5, "permutation is not a Bunch-Kaufman sequence");
    const int v = perm[a];
    if (v < 0 || v >= n) return h->fail(-5, "permutation entry out of range");
    const int b = pos[v];
    if (b < a) return h->fail(-5, "permutation is not a Bunch-Kaufman sequence");
    if (b != a) {
      std::swap(cur[a], cur[b]);
      pos[cur[a]] = a;
      pos[cur[b]] = b;
      steps.push_back(k);
      steps.push_back(a);
      steps.push_back(b);
    }
    if (two) {
      ipiv[k] = ipiv[k + 1] = -(b + 1);
      k += 2;
    } else {
      ipiv[k] = b + 1;
      k += 1;
    }
  }
  int rc = bbk::upload_and_apply(h, n, d_F, ldf, steps, true);
  if (rc != BBK_OK) return rc;
  e = hipMemcpyAsync(d_ipiv, ipiv.data(), sizeof(int) * n, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return h->hip_fail(e, "write ipiv");
  return BBK_OK;
}

int bbk_dsytrf_from_lapack(bbk_handle_t handle, int n, double* d_F, int64_t ldf,
                           const int* d_ipiv, int* d_perm, int* d_piv) {
  if (!handle) return -1;
  if (n <= 0) return -2;
  if (!d_F) return -3;
  if (ldf < n) return -4;
  if (!d_ipiv) return -5;
  if (!d_perm) return -6;
  if (!d_piv) return -7;
  Handle* h = reinterpret_cast<Handle*>(handle);
  std::vector<int> ipiv(n), perm(n), piv(n);
  hipError_t e = hipMemcpyAsync(ipiv.data(), d_ipiv, sizeof(int) * n, hipMemcpyDeviceToHost,
                                h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return h->hip_fail(e, "read ipiv");
  for (int i = 0; i < n; ++i) perm[i] = i;
  std::vector<int> steps;
  for (int k = 0; k < n;) {
    const int v = ipiv[k];
    int a, b;
    if (v > 0) {
      a = k;
      b = v - 1;
      piv[k] = 1;
    } else {
      if (v == 0 || k + 1 >= n || ipiv[k + 1] != v) return h->fail(-5, "invalid ipiv");
      a = k + 1;
      b = -v - 1;
      piv[k] = 2;
      piv[k + 1] = -2;
    }
    if (b < a || b >= n) return h->fail(-5, "invalid ipiv");
    if (b != a) {
      std::swap(perm[a], perm[b]);
      steps.push_back(k);
      steps.push_back(a);
      steps.push_back(b);
    }
    k += v > 0 ? 1 : 2;
  }
  int rc = bbk::upload_and_apply(h, n, d_F, ldf, steps, false);
  if (rc != BBK_OK) return rc;
  e = hipMemcpyAsync(d_perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_piv, piv.data(), sizeof(int) * n, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return h->hip_fail(e, "write pivots");
  return BBK_OK;
}

}  // extern "C"

// bbk_factor.hip -- gfx950 kernels of the blocked Bunch-Kaufman LDL^T
// factorization: the panel kernel (pivot search, symmetric interchanges,
// 1x1/2x2 scaling), the MFMA trailing update and the deferred
// interchange of earlier rows.
//
// Storage: A is row-major with leading dimension lda and only its upper
// triangle is used; row i of the upper triangle is column i of LAPACK's
// lower triangle, so "U row i" below is L(:, i)^T.  The panel's unscaled
// rows W = D L^T live in LDS while the panel runs and are flushed to a
// global (b+1) x ldw buffer for the trailing update.
#include <climits>

#include "bbk_kernels.hpp"
#include "bbk_update.hpp"

namespace bbk {

// ---------------------------------------------------------------------
// Deferred interchanges of the rows above the panel (the reference swaps
// them immediately, src/block_ldlt_decomposition.cc:472-491): the panel's
// interchange sequence (a_q, r_q), q < ns, is composed into one gather per
// row, row[dest_t] = row[src_t] over the ne = 2 ns touched columns, and
// applied to rows [p0, p1).  Every workgroup of this role composes the
// gather itself (ne <= 2 (kPanelMaxWidth + 1) entries).
template <typename T, int RPG>
__device__ __forceinline__ void swap_rows(T* __restrict__ A, int64_t lda,
                                          const int* __restrict__ swaps, int ns, int p0,
                                          int p1, int* cols, int* slot, int* content) {
  const int ne = 2 * ns;
  for (int t = threadIdx.x; t < ne; t += blockDim.x) cols[t] = swaps[t];
  __syncthreads();
  compose_swaps(cols, ns, slot, content);
  const int t0 = threadIdx.x, t1 = threadIdx.x + 256;
  const int d0 = t0 < ne ? cols[t0] : 0, s0 = t0 < ne ? content[slot[t0]] : 0;
  const int d1 = t1 < ne ? cols[t1] : 0, s1 = t1 < ne ? content[slot[t1]] : 0;
  T v0[RPG], v1[RPG];
#pragma unroll
  for (int u = 0; u < RPG; ++u) {
    const int row = p0 + u;
    const T* R = A + (int64_t)row * lda;
    v0[u] = (row < p1 && t0 < ne) ? R[s0] : T(0);
    v1[u] = (row < p1 && t1 < ne) ? R[s1] : T(0);
  }
  wait_vm();
  __syncthreads();  // every read of the rows precedes every write
#pragma unroll
  for (int u = 0; u < RPG; ++u) {
    const int row = p0 + u;
    T* R = A + (int64_t)row * lda;
    if (row < p1 && t0 < ne) R[d0] = v0[u];
    if (row < p1 && t1 < ne) R[d1] = v1[u];
  }
}

// One launch after each panel: workgroups [0, gemm_blocks) update trailing
// tiles (XCD-contiguous order), the rest apply the panel's interchanges to
// RPG rows each of the finished rows [rows_lo, s) above the panel.  The two
// regions are disjoint (rows >= e against rows < s), so both run
// concurrently.
template <typename T, int TM, int MINB, int KC, int RPG, bool SUPER = true>
__global__ __launch_bounds__(256, MINB) void bk_trailing_update_kernel(
    T* __restrict__ A, int64_t lda, const T* __restrict__ Wg, int64_t ldw,
    const int* __restrict__ state, const int* __restrict__ swaps, int n, int gemm_blocks,
    int rows_lo, int rows_hi) {
  constexpr int NE = 2 * (kPanelMaxWidth + 1);
  constexpr int kStage = 2 * KC * (TM + 4);  // elements of one operand's two buffers
  static_assert(3 * NE * sizeof(int) <= 2 * kStage * sizeof(T),
                "the swap role's index arrays share the operand buffers");
  // one LDS block serves both roles (a workgroup takes one), so its size
  // is the GEMM staging alone and eight workgroups fit per CU
  __shared__ __align__(16) T stage[2 * kStage];
  T(*Us)[KC * (TM + 4)] = reinterpret_cast<T(*)[KC * (TM + 4)]>(stage);
  T(*Ws)[KC * (TM + 4)] = reinterpret_cast<T(*)[KC * (TM + 4)]>(stage + kStage);
  int* cols = reinterpret_cast<int*>(stage);
  int* slot = cols + NE;
  int* content = slot + NE;
  const int s = state[PanelState::kStart];
  const int bw = state[PanelState::kWidth];
  int idx = blockIdx.x;
  if (idx < gemm_blocks) {
    // workgroups are dealt to the 8 XCDs round-robin; give XCD x a
    // contiguous run of tiles (neighbouring W and U strips share its L2)
    const int nb = gemm_blocks, per = (nb + 7) / 8, x = idx & 7, r = idx >> 3;
    const int full = nb - 8 * (per - 1);  // XCDs that get `per` tiles
    idx = x < full ? x * per + r : full * per + (x - full) * (per - 1) + r;
    update_tile<T, TM, KC, SUPER>(A, lda, Wg, ldw, s, bw, n, idx, Us, Ws);
  } else {
    const int ns = state[PanelState::kNumSwaps];
    const int p0 = rows_lo + (idx - gemm_blocks) * RPG;
    const int p1 = min(min(s, rows_hi), p0 + RPG);
    if (ns == 0 || p0 >= p1) return;
    swap_rows<T, RPG>(A, lda, swaps, ns, p0, p1, cols, slot, content);
  }
}

// The trailing update restricted to tile columns [ja, jb) of the trailing
// triangle (every tile I <= J of those columns), on a copy of that part of
// the matrix: C and U are bases shifted by their copies' origins, so both
// are indexed with the matrix's own row and column numbers.  Tiles in
// column-major triangle order, XCD-contiguous runs.
template <typename T, int TM, int MINB, int KC>
__global__ __launch_bounds__(256, MINB) void bk_slab_update_kernel(
    T* __restrict__ C, int64_t ldc, const T* __restrict__ U, int64_t ldu,
    const T* __restrict__ Wg, int64_t ldw, const int* __restrict__ state, int n, int ja,
    int count) {
  constexpr int kStage = 2 * KC * (TM + 4);
  __shared__ __align__(16) T stage[2 * kStage];
  T(*Us)[KC * (TM + 4)] = reinterpret_cast<T(*)[KC * (TM + 4)]>(stage);
  T(*Ws)[KC * (TM + 4)] = reinterpret_cast<T(*)[KC * (TM + 4)]>(stage + kStage);
  const int s = state[PanelState::kStart];
  const int bw = state[PanelState::kWidth];
  if (bw <= 0 || s + bw >= n) return;
  const int nb = count, per = (nb + 7) / 8, x = blockIdx.x & 7, r = blockIdx.x >> 3;
  const int full = nb - 8 * (per - 1);
  const int idx = x < full ? x * per + r : full * per + (x - full) * (per - 1) + r;
  if (idx >= count) return;
  int I, J;
  tri_tile(ja * (ja + 1) / 2 + idx, &I, &J);
  update_tile_at<T, TM, KC>(C, ldc, U, ldu, Wg, ldw, s, bw, n, I, J, Us, Ws);
}

// Tile shape (measured on MI355X at n = 16384, K = 128, inside the
// factorization): binary64 64 x 64 tiles with eight workgroups resident per
// CU, so one workgroup's C traffic overlaps another's MFMAs (42.6 ms for all
// updates; C added after the K loop at five per CU 47.5 ms, operands two
// chunks ahead at seven per CU 44.8 ms, K chunks of 12 at six per CU
// 45.5 ms, K chunks of 16 at four per CU 51.0 ms); binary32 128 x 128 tiles, five per
// CU by registers (27.3 ms against 27.8 ms for 64 x 64 and 28.4 ms with K
// chunks of 16).  K chunks of 8; XCD-contiguous supertile order.
template <typename T>
struct UpdateCfg {
  static constexpr int kTM = sizeof(T) == 4 ? 128 : 64;
  static constexpr int kMinBlocks = sizeof(T) == 4 ? 4 : 8;
  static constexpr int kKC = 8;
};

__global__ void bk_init_kernel(int n, int* __restrict__ perm,
                               int* __restrict__ piv, int* __restrict__ state,
                               unsigned* __restrict__ status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    perm[i] = i;
    piv[i] = 0;
  }
  if (i < PanelState::kInts) state[i] = 0;
  if (status && i == 0) {
    status[kStatusInfo] = kInfoNone;
    status[kStatusNonFinite] = 0u;
  }
}

__global__ void bk_zero_u32_kernel(unsigned* __restrict__ p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0u;
}

// After a factorization: the first singular pivot (LAPACK's info: column k
// + 1 of a 1x1 pivot with |d| <= DBL_EPSILON, whose L column the scaling
// set to zero, or of a 2x2 block with a zero determinant, evaluated in
// binary64 without contraction as the scaling does) and whether D holds a
// NaN or an infinity, into the status words.  One workgroup; O(n) reads.
template <typename T>
__global__ __launch_bounds__(256) void bk_factor_info_kernel(int n, const T* __restrict__ A,
                                                             int64_t lda,
                                                             const int* __restrict__ piv,
                                                             unsigned* __restrict__ status) {
#pragma clang fp contract(off)
  __shared__ unsigned smin[4];
  __shared__ unsigned snf[4];
  unsigned first = kInfoNone, nf = 0u;
  for (int k = threadIdx.x; k < n; k += 256) {
    const int t = piv[k];
    const double d = static_cast<double>(A[(int64_t)k * lda + k]);
    if (!isfinite(d)) nf = 1u;
    bool singular = false;
    if (t == 2 && k + 1 < n) {
      const double b = static_cast<double>(A[(int64_t)k * lda + k + 1]);
      const double c = static_cast<double>(A[(int64_t)(k + 1) * lda + k + 1]);
      if (!isfinite(b)) nf = 1u;
      singular = d * c - b * b == 0.0;
    } else if (t == 1) {
      singular = fabs(d) <= kDblEps;
    }
    if (singular) first = min(first, static_cast<unsigned>(k + 1));
  }
  // block reduction: minimum index, or of the flags
  for (int off = 32; off > 0; off >>= 1) {
    first = min(first, static_cast<unsigned>(__shfl_xor(static_cast<int>(first), off, 64)));
    nf |= static_cast<unsigned>(__shfl_xor(static_cast<int>(nf), off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    smin[threadIdx.x >> 6] = first;
    snf[threadIdx.x >> 6] = nf;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned m = smin[0], f = snf[0];
    for (int w = 1; w < 4; ++w) {
      m = min(m, smin[w]);
      f |= snf[w];
    }
    status[kStatusInfo] = m;
    status[kStatusNonFinite] = f;
  }
}

template <typename T>
void launch_factor_info(int n, const T* A, int64_t lda, const int* piv, unsigned* status,
                        hipStream_t stream) {
  hipLaunchKernelGGL(bk_factor_info_kernel<T>, dim3(1), dim3(256), 0, stream, n, A, lda, piv,
                     status);
}
template void launch_factor_info<double>(int, const double*, int64_t, const int*, unsigned*,
                                         hipStream_t);
template void launch_factor_info<float>(int, const float*, int64_t, const int*, unsigned*,
                                        hipStream_t);

// ---------------------------------------------------------------------
// Launchers.
template <typename T>
void launch_trailing_update(T* A, int64_t lda, const T* Wg, int64_t ldw,
                            const int* state, const int* swaps, int n, int m_bound,
                            int rows_bound, hipStream_t stream, int rows_lo) {
  constexpr int TM = UpdateCfg<T>::kTM;
  constexpr int RPG = 8;
  const int tiles = m_bound > 0 ? (m_bound + TM - 1) / TM : 0;
  const int gemm_blocks = tiles * (tiles + 1) / 2;
  const int swap_blocks = rows_bound > rows_lo ? (rows_bound - rows_lo + RPG - 1) / RPG : 0;
  if (gemm_blocks + swap_blocks == 0) return;
  hipLaunchKernelGGL((bk_trailing_update_kernel<T, TM, UpdateCfg<T>::kMinBlocks,
                                                 UpdateCfg<T>::kKC, RPG>),
                     dim3(gemm_blocks + swap_blocks), dim3(256), 0, stream, A, lda, Wg, ldw,
                     state, swaps, n, gemm_blocks, rows_lo, rows_bound);
}

template <typename T>
int slab_tile_width() {
  return UpdateCfg<T>::kTM;
}

template <typename T>
void launch_slab_update(T* C, int64_t ldc, const T* U, int64_t ldu, const T* Wg, int64_t ldw,
                        const int* state, int n, int ja, int jb, hipStream_t stream) {
  constexpr int TM = UpdateCfg<T>::kTM;
  const int count = jb * (jb + 1) / 2 - ja * (ja + 1) / 2;
  if (count <= 0) return;
  hipLaunchKernelGGL((bk_slab_update_kernel<T, TM, UpdateCfg<T>::kMinBlocks, UpdateCfg<T>::kKC>),
                     dim3(count), dim3(256), 0, stream, C, ldc, U, ldu, Wg, ldw, state, n, ja,
                     count);
}

void launch_init(int n, int* perm, int* piv, int* state, unsigned* status, hipStream_t stream) {
  const int blocks = (n + 255) / 256;
  hipLaunchKernelGGL(bk_init_kernel, dim3(blocks > 0 ? blocks : 1), dim3(256),
                     0, stream, n, perm, piv, state, status);
}

void launch_zero_u32(unsigned* p, int count, hipStream_t stream) {
  hipLaunchKernelGGL(bk_zero_u32_kernel, dim3((count + 255) / 256), dim3(256),
                     0, stream, p, count);
}

template void launch_trailing_update<double>(double*, int64_t, const double*, int64_t,
                                             const int*, const int*, int, int, int,
                                             hipStream_t, int);
template void launch_trailing_update<float>(float*, int64_t, const float*, int64_t,
                                            const int*, const int*, int, int, int,
                                            hipStream_t, int);
template int slab_tile_width<double>();
template void launch_slab_update<double>(double*, int64_t, const double*, int64_t, const double*,
                                         int64_t, const int*, int, int, int, hipStream_t);

}  // namespace bbk

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

namespace bbk {

// ---------------------------------------------------------------------
// Pivot arithmetic kept free of FMA contraction so the decisions and the
// 2x2 inverse round exactly like the reference
// (src/block_ldlt_decomposition.cc:369-390, 658-677).
__device__ __forceinline__ int select_pivot_nc(double akk, double omega,
                                            double sigma, double arr) {
#pragma clang fp contract(off)
  if (akk * sigma - kAlpha * omega * omega >= 0.0) return 3;
  if (arr - kAlpha * sigma >= 0.0) return -1;
  return 2;
}

__device__ __forceinline__ void inverse_2x2_nc(double a, double b, double c,
                                            double* itl, double* ibr,
                                            double* ioff) {
#pragma clang fp contract(off)
  const double det = a * c - b * b;
  *itl = c / det;
  *ibr = a / det;
  *ioff = -b / det;
}

__device__ __forceinline__ void apply_2x2_nc(double itl, double ibr,
                                             double ioff, double w0,
                                             double w1, double* u0,
                                             double* u1) {
#pragma clang fp contract(off)
  *u0 = itl * w0 + ioff * w1;
  *u1 = ioff * w0 + ibr * w1;
}

__device__ __forceinline__ double safe_den(double d) {
  // SafeDenominator, src/block_ldlt_decomposition.cc:21-26
  return fabs(d) <= kDblEps ? 0.0 : d;
}

// ---------------------------------------------------------------------
// Panel kernel.  Launched cooperatively with G workgroups of
// kPanelThreads threads; workgroup g owns columns
// [s + g*NT, s + (g+1)*NT) of the active region [s, n).  Per pivot:
//   A  row k of W = A(k,:) - L(k,panel) W(panel,:)   -> argmax, barrier
//   B  candidate row r (if the BK test needs it)     -> max, barrier
//   C  interchange + scale                           -> barrier
template <typename T>
__global__ __launch_bounds__(kPanelThreads) void bk_panel_kernel(
    T* __restrict__ A, PanelArgs p) {
  constexpr int NT = kPanelThreads;
  constexpr int BM = PanelCfg<T>::kMaxWidth;
  typedef PanelExchange X_;
  __shared__ T Wl[(BM + 1) * NT];
  __shared__ T Lk[BM + 1];
  __shared__ T Lk1[BM + 1];
  __shared__ T Lr[BM + 1];
  __shared__ double red_v[NT / 64];
  __shared__ int red_i[NT / 64];

  const int n = p.n;
  const int64_t lda = p.lda;
  const int s = p.state[PanelState::kNextStart];
  if (s >= n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      p.state[PanelState::kStart] = s;
      p.state[PanelState::kWidth] = 0;
      p.state[PanelState::kNumSwaps] = 0;
    }
    return;
  }
  const unsigned G = gridDim.x;
  const int tid = threadIdx.x;
  const int j = s + blockIdx.x * NT + tid;
  const bool own = j < n;
  const int bt = min(p.b, n - s);
  T* Wg = reinterpret_cast<T*>(p.Wg);
  double* X = p.x;
  unsigned nbar = 0;
  int kk = 0, nswap = 0;

  while (kk < bt) {
    const int k = s + kk;
    // ---- phase A: W(k, j) = A(k, j) - sum_q L(k, s+q) W(q, j) ----------
    for (int q = tid; q < kk; q += NT) {
      Lk[q] = A[(int64_t)(s + q) * lda + k];
      Lk1[q] = (k + 1 < n) ? A[(int64_t)(s + q) * lda + k + 1] : T(0);
    }
    __syncthreads();
    const bool active = own && j >= k;
    T w = T(0), arow = T(0);
    if (active) {
      arow = A[(int64_t)k * lda + j];
      T upd = T(0);
      for (int q = 0; q < kk; ++q) upd += Lk[q] * Wl[q * NT + tid];
      w = arow - upd;
      Wl[kk * NT + tid] = w;
    }
    double v = 0.0;
    int vi = INT_MAX;
    if (active && j > k) {
      const double a = fabs(static_cast<double>(w));
      if (a > 0.0) {  // NaN never selected (strict '>' scan)
        v = a;
        vi = j;
      }
    }
    block_argmax(v, vi, red_v, red_i);
    if (tid == 0) {
      X[X_::kOmega + blockIdx.x] = v;
      p.xi[blockIdx.x] = vi;
    }
    if (j == k) {
      X[X_::kScalars + X_::sWkk] = static_cast<double>(w);
      X[X_::kScalars + X_::sAkkOld] = static_cast<double>(arow);
    }
    grid_barrier(p.bar, (++nbar) * G);

    double omega = 0.0;
    int r = INT_MAX;
    for (unsigned g2 = tid; g2 < G; g2 += NT)
      argmax_combine(omega, r, X[X_::kOmega + g2], p.xi[g2]);
    block_argmax(omega, r, red_v, red_i);
    const double wkk = X[X_::kScalars + X_::sWkk];
    const double akk = fabs(wkk);

    int kind;
    T c = T(0);
    if (akk >= kAlpha * omega || omega == 0.0) {
      kind = 1;  // src/block_ldlt_decomposition.cc:92
    } else {
      // ---- phase B: candidate row r, sigma = max_{i>=k, i!=r} |W(i,r)| --
      for (int q = tid; q < kk; q += NT) Lr[q] = A[(int64_t)(s + q) * lda + r];
      __syncthreads();
      if (active) {
        const T a = (j < r) ? A[(int64_t)j * lda + r] : A[(int64_t)r * lda + j];
        T upd = T(0);
        for (int q = 0; q < kk; ++q) upd += Lr[q] * Wl[q * NT + tid];
        c = a - upd;
      }
      double sv = 0.0;
      if (active && j != r) {
        const double a = fabs(static_cast<double>(c));
        sv = (a >= 0.0) ? a : 0.0;  // NaN ignored, as std::max does
      }
      sv = block_max_nan_ignoring(sv, red_v);
      if (tid == 0) X[X_::kSigma + blockIdx.x] = sv;
      if (j == r) {
        X[X_::kScalars + X_::sCrr] = static_cast<double>(c);
        X[X_::kScalars + X_::sWkr] = static_cast<double>(Wl[kk * NT + tid]);
        for (int q = 0; q < kk; ++q)
          X[X_::kColR + q] = static_cast<double>(Wl[q * NT + tid]);
      }
      if (j == k) {
        X[X_::kScalars + X_::sCk] = static_cast<double>(c);
        for (int q = 0; q < kk; ++q)
          X[X_::kColK + q] = static_cast<double>(Wl[q * NT + tid]);
      }
      if (j == k + 1) {
        X[X_::kScalars + X_::sCk1] = static_cast<double>(c);
        X[X_::kScalars + X_::sWk1] = static_cast<double>(Wl[kk * NT + tid]);
        X[X_::kScalars + X_::sAk1Old] =
            static_cast<double>(A[(int64_t)(k + 1) * lda + k + 1]);
        for (int q = 0; q < kk; ++q)
          X[X_::kColK1 + q] = static_cast<double>(Wl[q * NT + tid]);
      }
      grid_barrier(p.bar, (++nbar) * G);
      double sigma = 0.0;
      for (unsigned g2 = tid; g2 < G; g2 += NT) {
        const double o = X[X_::kSigma + g2];
        sigma = (sigma < o) ? o : sigma;
      }
      sigma = block_max_nan_ignoring(sigma, red_v);
      const double crr = X[X_::kScalars + X_::sCrr];
      kind = select_pivot_nc(akk, omega, sigma, fabs(crr));
    }

    // ---- phase C: interchange and scale --------------------------------
    int size;
    if (kind == 1 || kind == 3) {
      // ScaleOneByOnePivot, src/block_ldlt_decomposition.cc:633-656
      const double den = safe_den(wkk);
      if (own && j > k)
        A[(int64_t)k * lda + j] =
            den == 0.0 ? T(0) : static_cast<T>(static_cast<double>(w) / den);
      if (j == k) {
        A[(int64_t)k * lda + k] = w;
        p.piv[k] = 1;
      }
      size = 1;
    } else if (kind == -1) {
      // symmetric interchange k <-> r (src/block_ldlt_decomposition.cc:414-498)
      if (j == k) {
        for (int q = 0; q < kk; ++q) {
          Wl[q * NT + tid] = static_cast<T>(X[X_::kColR + q]);
          A[(int64_t)(s + q) * lda + k] = Lr[q];
        }
      }
      if (j == r) {
        for (int q = 0; q < kk; ++q) {
          Wl[q * NT + tid] = static_cast<T>(X[X_::kColK + q]);
          A[(int64_t)(s + q) * lda + r] = Lk[q];
        }
      }
      T nw = c;
      if (j == k)
        nw = static_cast<T>(X[X_::kScalars + X_::sCrr]);
      else if (j == r)
        nw = static_cast<T>(X[X_::kScalars + X_::sCk]);
      if (active) Wl[kk * NT + tid] = nw;
      // trailing matrix: new row/column r take the old row k
      if (own && j > r)
        A[(int64_t)r * lda + j] = arow;
      else if (own && j > k && j < r)
        A[(int64_t)j * lda + r] = arow;
      if (j == r)
        A[(int64_t)r * lda + r] = static_cast<T>(X[X_::kScalars + X_::sAkkOld]);
      const double d = X[X_::kScalars + X_::sCrr];
      const double den = safe_den(d);
      if (own && j > k)
        A[(int64_t)k * lda + j] =
            den == 0.0 ? T(0) : static_cast<T>(static_cast<double>(nw) / den);
      if (j == k) {
        A[(int64_t)k * lda + k] = nw;
        p.piv[k] = 1;
      }
      if (blockIdx.x == 0 && tid == 0) {
        const int t = p.perm[k];
        p.perm[k] = p.perm[r];
        p.perm[r] = t;
        p.swaps[2 * nswap] = k;
        p.swaps[2 * nswap + 1] = r;
      }
      ++nswap;
      size = 1;
    } else {
      // 2x2 pivot at (k, k+1) after the interchange k+1 <-> r
      // (src/block_ldlt_decomposition.cc:500-587, 658-704)
      const int k1 = k + 1;
      if (r != k1) {
        if (j == k1) {
          for (int q = 0; q < kk; ++q) {
            Wl[q * NT + tid] = static_cast<T>(X[X_::kColR + q]);
            A[(int64_t)(s + q) * lda + k1] = Lr[q];
          }
        }
        if (j == r) {
          for (int q = 0; q < kk; ++q) {
            Wl[q * NT + tid] = static_cast<T>(X[X_::kColK1 + q]);
            A[(int64_t)(s + q) * lda + r] = Lk1[q];
          }
        }
        const T ak1 = (own && j > k1) ? A[(int64_t)k1 * lda + j] : T(0);
        if (own && j > r)
          A[(int64_t)r * lda + j] = ak1;
        else if (own && j > k1 && j < r)
          A[(int64_t)j * lda + r] = ak1;
        if (j == r)
          A[(int64_t)r * lda + r] =
              static_cast<T>(X[X_::kScalars + X_::sAk1Old]);
        if (blockIdx.x == 0 && tid == 0) {
          const int t = p.perm[k1];
          p.perm[k1] = p.perm[r];
          p.perm[r] = t;
          p.swaps[2 * nswap] = k1;
          p.swaps[2 * nswap + 1] = r;
        }
        ++nswap;
      }
      T w0 = active ? Wl[kk * NT + tid] : T(0);
      T w1 = c;
      if (j == k1) {
        w0 = static_cast<T>(X[X_::kScalars + X_::sWkr]);
        w1 = static_cast<T>(X[X_::kScalars + X_::sCrr]);
      } else if (j == r) {
        w0 = static_cast<T>(X[X_::kScalars + X_::sWk1]);
        w1 = static_cast<T>(X[X_::kScalars + X_::sCk1]);
      }
      if (active) {
        Wl[kk * NT + tid] = w0;
        Wl[(kk + 1) * NT + tid] = w1;
      }
      const double a = X[X_::kScalars + X_::sWkk];
      const double bb = X[X_::kScalars + X_::sWkr];
      const double cc = X[X_::kScalars + X_::sCrr];
      double itl, ibr, ioff;
      inverse_2x2_nc(a, bb, cc, &itl, &ibr, &ioff);
      if (own && j >= k + 2) {
        double u0, u1;
        apply_2x2_nc(itl, ibr, ioff, static_cast<double>(w0),
                     static_cast<double>(w1), &u0, &u1);
        A[(int64_t)k * lda + j] = static_cast<T>(u0);
        A[(int64_t)k1 * lda + j] = static_cast<T>(u1);
      }
      if (j == k) {
        A[(int64_t)k * lda + k] = static_cast<T>(a);
        p.piv[k] = 2;
      }
      if (j == k1) {
        A[(int64_t)k * lda + k1] = static_cast<T>(bb);
        A[(int64_t)k1 * lda + k1] = static_cast<T>(cc);
        p.piv[k1] = -2;
      }
      size = 2;
    }
    grid_barrier(p.bar, (++nbar) * G);
    kk += size;
  }

  // W rows of this panel for the trailing update.
  if (own)
    for (int q = 0; q < kk; ++q) Wg[(int64_t)q * p.ldw + j] = Wl[q * NT + tid];
  if (blockIdx.x == 0 && tid == 0) {
    p.state[PanelState::kStart] = s;
    p.state[PanelState::kWidth] = kk;
    p.state[PanelState::kNumSwaps] = nswap;
    p.state[PanelState::kNextStart] = s + kk;
  }
}

// ---------------------------------------------------------------------
// Trailing update C(i,j) -= sum_q U(s+q, i) W(q, j) over the upper
// triangle of [e, n)^2, e = s + width (UpdateTrailingBlocks +
// kernels::UpdateTrailingBlock, src/block_ldlt_decomposition.cc:706-744,
// src/kernels.cc:67-91).  128x128 output tile per workgroup, four
// wavefronts of 64x64, 16x16x4 MFMA, K staged through LDS in chunks.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output element v of a lane sits in row (lane / 16) + 4 v (measured)
  static __device__ __forceinline__ int row(int lane, int v) {
    return (lane >> 4) + 4 * v;
  }
};
template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // output element v of a lane sits in row 4 (lane / 16) + v (measured)
  static __device__ __forceinline__ int row(int lane, int v) {
    return 4 * (lane >> 4) + v;
  }
};

__device__ __forceinline__ void tri_tile(int idx, int* I, int* J) {
  int jj = static_cast<int>((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
  while (jj * (jj + 1) / 2 > idx) --jj;
  while ((jj + 1) * (jj + 2) / 2 <= idx) ++jj;
  *J = jj;
  *I = idx - jj * (jj + 1) / 2;
}

template <typename T>
__global__ __launch_bounds__(256) void bk_trailing_update_kernel(
    T* __restrict__ A, int64_t lda, const T* __restrict__ Wg, int64_t ldw,
    const int* __restrict__ state, int n) {
  constexpr int TM = 128;
  constexpr int KC = 16;
  constexpr int LDS_LD = TM + 4;
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T Us[KC * LDS_LD];
  __shared__ T Ws[KC * LDS_LD];

  const int s = state[PanelState::kStart];
  const int bw = state[PanelState::kWidth];
  const int e = s + bw;
  const int m = n - e;
  if (m <= 0 || bw <= 0) return;
  const int tiles = (m + TM - 1) / TM;
  int I, J;
  tri_tile(blockIdx.x, &I, &J);
  if (J >= tiles) return;
  const int i0 = e + I * TM, j0 = e + J * TM;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  acc_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

  for (int k0 = 0; k0 < bw; k0 += KC) {
    for (int t = threadIdx.x; t < KC * TM; t += 256) {
      const int q = t / TM, cidx = t % TM;
      const int kq = k0 + q;
      T u = T(0), wv = T(0);
      if (kq < bw) {
        if (i0 + cidx < n) u = A[(int64_t)(s + kq) * lda + i0 + cidx];
        if (j0 + cidx < n) wv = Wg[(int64_t)kq * ldw + j0 + cidx];
      }
      Us[q * LDS_LD + cidx] = u;
      Ws[q * LDS_LD + cidx] = wv;
    }
    __syncthreads();
#pragma unroll
    for (int k4 = 0; k4 < KC; k4 += 4) {
      T a[4], b[4];
      const int kr = k4 + (lane >> 4);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        a[t] = Us[kr * LDS_LD + wr * 64 + t * 16 + (lane & 15)];
        b[t] = Ws[kr * LDS_LD + wc * 64 + t * 16 + (lane & 15)];
      }
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
          acc[ti][tj] = Mfma<T>::mma(a[ti], b[tj], acc[ti][tj]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wr * 64 + ti * 16 + Mfma<T>::row(lane, v);
        const int jj = j0 + wc * 64 + tj * 16 + (lane & 15);
        if (i < n && jj < n && jj >= i) A[(int64_t)i * lda + jj] -= acc[ti][tj][v];
      }
}

// ---------------------------------------------------------------------
// Deferred interchanges of the rows above the panel: the panel's
// interchange sequence composed into one gather per row
// (the reference swaps these immediately, src/block_ldlt_decomposition.cc:472-491).
__global__ __launch_bounds__(256) void bk_swap_plan_kernel(
    const int* __restrict__ state, const int* __restrict__ swaps,
    int* __restrict__ plan) {
  constexpr int NE = 2 * (kPanelMaxWidth + 1);
  __shared__ int cols[NE];
  __shared__ int slot[NE];
  __shared__ int content[NE];
  const int ns = state[PanelState::kNumSwaps];
  const int ne = 2 * ns;
  for (int t = threadIdx.x; t < ne; t += blockDim.x) cols[t] = swaps[t];
  __syncthreads();
  for (int t = threadIdx.x; t < ne; t += blockDim.x) {
    const int cval = cols[t];
    int f = t;
    for (int u = 0; u < t; ++u)
      if (cols[u] == cval) {
        f = u;
        break;
      }
    slot[t] = f;
    content[t] = cval;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < ns; ++q) {
      const int sa = slot[2 * q], sb = slot[2 * q + 1];
      const int tmp = content[sa];
      content[sa] = content[sb];
      content[sb] = tmp;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ne; t += blockDim.x) {
    plan[t] = cols[t];                 // destination column
    plan[NE + t] = content[slot[t]];   // source column
  }
}

template <typename T>
__global__ __launch_bounds__(256) void bk_swap_apply_kernel(
    T* __restrict__ A, int64_t lda, const int* __restrict__ state,
    const int* __restrict__ plan, int rows_per_group) {
  constexpr int NE = 2 * (kPanelMaxWidth + 1);
  const int s = state[PanelState::kStart];
  const int ne = 2 * state[PanelState::kNumSwaps];
  if (ne == 0) return;
  const int p0 = blockIdx.x * rows_per_group;
  if (p0 >= s) return;
  const int p1 = min(s, p0 + rows_per_group);
  const int t0 = threadIdx.x, t1 = threadIdx.x + 256;
  const int d0 = t0 < ne ? plan[t0] : 0, s0 = t0 < ne ? plan[NE + t0] : 0;
  const int d1 = t1 < ne ? plan[t1] : 0, s1 = t1 < ne ? plan[NE + t1] : 0;
  for (int row = p0; row < p1; ++row) {
    T* R = A + (int64_t)row * lda;
    const T v0 = t0 < ne ? R[s0] : T(0);
    const T v1 = t1 < ne ? R[s1] : T(0);
    __syncthreads();
    if (t0 < ne) R[d0] = v0;
    if (t1 < ne) R[d1] = v1;
    __syncthreads();
  }
}

__global__ void bk_init_kernel(int n, int* __restrict__ perm,
                               int* __restrict__ piv, int* __restrict__ state) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    perm[i] = i;
    piv[i] = 0;
  }
  if (i < PanelState::kInts) state[i] = 0;
}

__global__ void bk_zero_u32_kernel(unsigned* __restrict__ p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0u;
}

// ---------------------------------------------------------------------
// Launchers.
template <typename T>
hipError_t launch_panel(T* A, const PanelArgs& args, int groups,
                        hipStream_t stream) {
  T* a = A;
  PanelArgs pa = args;
  void* kargs[] = {&a, &pa};
  return hipLaunchCooperativeKernel(
      reinterpret_cast<const void*>(&bk_panel_kernel<T>), dim3(groups),
      dim3(kPanelThreads), kargs, 0, stream);
}

template <typename T>
hipError_t panel_max_groups(int* groups) {
  int per_cu = 0, dev = 0, cus = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(&bk_panel_kernel<T>),
      kPanelThreads, 0);
  if (e != hipSuccess) return e;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  *groups = per_cu * cus;
  return hipSuccess;
}

template <typename T>
void launch_trailing_update(T* A, int64_t lda, const T* Wg, int64_t ldw,
                            const int* state, int n, int m_bound,
                            hipStream_t stream) {
  if (m_bound <= 0) return;
  const int tiles = (m_bound + 127) / 128;
  const int blocks = tiles * (tiles + 1) / 2;
  hipLaunchKernelGGL(bk_trailing_update_kernel<T>, dim3(blocks), dim3(256), 0,
                     stream, A, lda, Wg, ldw, state, n);
}

template <typename T>
void launch_swap_rows(T* A, int64_t lda, const int* state, const int* swaps,
                      int* plan, int rows_bound, hipStream_t stream) {
  if (rows_bound <= 0) return;
  hipLaunchKernelGGL(bk_swap_plan_kernel, dim3(1), dim3(256), 0, stream, state,
                     swaps, plan);
  const int rpg = 4;
  const int groups = (rows_bound + rpg - 1) / rpg;
  hipLaunchKernelGGL(bk_swap_apply_kernel<T>, dim3(groups), dim3(256), 0,
                     stream, A, lda, state, plan, rpg);
}

void launch_init(int n, int* perm, int* piv, int* state, hipStream_t stream) {
  const int blocks = (n + 255) / 256;
  hipLaunchKernelGGL(bk_init_kernel, dim3(blocks > 0 ? blocks : 1), dim3(256),
                     0, stream, n, perm, piv, state);
}

void launch_zero_u32(unsigned* p, int count, hipStream_t stream) {
  hipLaunchKernelGGL(bk_zero_u32_kernel, dim3((count + 255) / 256), dim3(256),
                     0, stream, p, count);
}

template hipError_t launch_panel<double>(double*, const PanelArgs&, int,
                                         hipStream_t);
template hipError_t launch_panel<float>(float*, const PanelArgs&, int,
                                        hipStream_t);
template hipError_t panel_max_groups<double>(int*);
template hipError_t panel_max_groups<float>(int*);
template void launch_trailing_update<double>(double*, int64_t, const double*,
                                             int64_t, const int*, int, int,
                                             hipStream_t);
template void launch_trailing_update<float>(float*, int64_t, const float*,
                                            int64_t, const int*, int, int,
                                            hipStream_t);
template void launch_swap_rows<double>(double*, int64_t, const int*,
                                       const int*, int*, int, hipStream_t);
template void launch_swap_rows<float>(float*, int64_t, const int*, const int*,
                                      int*, int, hipStream_t);

}  // namespace bbk

// bbk_panel.hip -- the Bunch-Kaufman panel kernel for gfx950.
//
// One cooperative launch factors one panel of up to b columns starting at
// s (read from device state).  Workgroup g owns the active columns
// [s + 256 g, s + 256 g + 256) of rows [s, n); thread t owns one column and
// keeps that column of the panel's unscaled rows W = D L^T in LDS.
//
// Per pivot k (local index kk) the workgroups meet at most twice:
//   exchange 1  after W(k, :) = A(k, :) - L(k, panel) W(panel, :) and the
//               local argmax; carries (omega_g, r_g), W(panel, r_g), the
//               W columns k, k+1, k+2 and a few scalars;
//   exchange 2  (only when the Bunch-Kaufman test needs the candidate row r)
//               after the candidate row and its local max sigma_g; carries
//               sigma_g and the candidate values at k, k+1, k+2, r.
// Every datum travels as a 16-byte record {tag, aux, value} written and
// read with device-coherent (sc1) vector accesses; a reader accepts a
// record only when its tag equals the exchange's number, so an exchange is
// one store per writer and one polling read per reader, with no atomics.
// Two record areas alternate by exchange parity.  Everything the next
// pivot needs (its L column, its A row) is rebuilt locally from published
// values, so no third synchronization per pivot is needed.
//
// Arithmetic: pivot tests, 2x2 inverse and L = D^-1 W in binary64 without
// contraction (src/block_ldlt_decomposition.cc:21-26,369-390,633-704).
#include <climits>

#include "bbk_kernels.hpp"

namespace bbk {

#if defined(BBK_PANEL_PROFILE)
// Per-phase clock totals of groups 0 (slots 0-7) and G-1 (slots 8-15).
__device__ unsigned long long g_panel_profile[16];
#endif

// Stores that stay in the XCD's L2 (scope bit sc0).  A device-coherent (sc1)
// load from another CU of the same XCD sees them (measured on MI355X:
// 0.66 us ping-pong against 1.13 us with sc1 stores; across XCDs such a
// store is not seen until written back, so these are used only when every
// working group of the launch sits on one XCD).
__device__ __forceinline__ void xput_l2(XRec* p, int tag, int aux, double v) {
  const long long b = __double_as_longlong(v);
  v4i d = {tag, aux, static_cast<int>(b & 0xffffffffll), static_cast<int>(b >> 32)};
  asm volatile("global_store_dwordx4 %0, %1, off sc0" ::"v"(p), "v"(d) : "memory");
}
template <typename T>
__device__ __forceinline__ void st_l2(T* p, T v);
template <>
__device__ __forceinline__ void st_l2<double>(double* p, double v) {
  asm volatile("global_store_dwordx2 %0, %1, off sc0" ::"v"(p), "v"(v) : "memory");
}
template <>
__device__ __forceinline__ void st_l2<float>(float* p, float v) {
  asm volatile("global_store_dword %0, %1, off sc0" ::"v"(p), "v"(v) : "memory");
}

// Accesses to A inside the panel kernel are device-coherent (sc1) so that
// no cache-wide writeback/invalidate is needed around an exchange: a
// writer only waits for its stores to complete before posting records.
template <typename T>
__device__ __forceinline__ T ldc(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void stc(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// sum_q L[q] * W(q, own column), q < kk; four accumulators keep several
// LDS reads in flight.
template <typename T, int NT>
__device__ __forceinline__ T panel_dot(const T* L, const T* Wl, int tid, int kk) {
  T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
  int q = 0;
#pragma unroll 2
  for (; q + 4 <= kk; q += 4) {
    a0 += L[q] * Wl[q * NT + tid];
    a1 += L[q + 1] * Wl[(q + 1) * NT + tid];
    a2 += L[q + 2] * Wl[(q + 2) * NT + tid];
    a3 += L[q + 3] * Wl[(q + 3) * NT + tid];
  }
  for (; q < kk; ++q) a0 += L[q] * Wl[q * NT + tid];
  return (a0 + a1) + (a2 + a3);
}

// Panel-row scaling table: row q of L^T = D^-1 W.
//   type 1 : U(q) = W(q) / c0            (c0 = SafeDenominator(d), 0 -> 0)
//   type 2 : U(q) = c0 W(q) + c1 W(q+1)  (c0 = itl, c1 = ioff)
//   type -2: U(q) = c0 W(q-1) + c1 W(q)  (c0 = ioff, c1 = ibr)
__device__ __forceinline__ double ucol(const int* dt, const double* c0, const double* c1,
                                       const double* col, int q) {
  const int ty = dt[q];
  if (ty == 1) return div_or_zero(col[q], c0[q]);
  double u0, u1;
  if (ty == 2) {
    apply_2x2_nc(c0[q], 0.0, c1[q], col[q], col[q + 1], &u0, &u1);
    return u0;
  }
  apply_2x2_nc(0.0, c1[q], c0[q], col[q - 1], col[q], &u0, &u1);
  return u1;
}

enum PubId {
  kPubWk = 0, kPubAkk = 1, kPubWk1 = 2, kPubAkk1 = 3, kPubWk2 = 4,
  kPubCr = 8, kPubCk = 9, kPubCk1 = 10, kPubA11 = 11, kPubCk2 = 12, kPubA12 = 13,
  kNumPub = 16
};

// L(r, s + q) from W(q, r) (the candidate's column) and row q's scaling
// record; `partner` is W(q + 1, r) for a 2x2 start, W(q - 1, r) for its
// second row.  Same operations as ucol().
__device__ __forceinline__ double ucol_pair(int ty, double c0, double c1, double self,
                                            double partner) {
  if (ty == 1) return div_or_zero(self, c0);
  double u0, u1;
  if (ty == 2) {
    apply_2x2_nc(c0, 0.0, c1, self, partner, &u0, &u1);
    return u0;
  }
  apply_2x2_nc(0.0, c1, c0, partner, self, &u0, &u1);
  return u1;
}

template <typename T>
__global__ __launch_bounds__(kPanelThreads) void bk_panel_kernel(T* __restrict__ A, PanelArgs p) {
  constexpr int NT = kPanelThreads;
  constexpr int BM = PanelCfg<T>::kMaxWidth;
  __shared__ T Wl[(BM + 1) * NT];
  __shared__ T Lbuf[2][BM + 2];  // L(k, panel) of this pivot / of the next
  __shared__ T Lr[BM + 2];
  __shared__ double colR[BM + 2], colK[BM + 2], colK1[BM + 2], colK2[BM + 2];
  __shared__ int dt[BM + 2];
  __shared__ double dc0[BM + 2], dc1[BM + 2];
  __shared__ double xval[kMaxPanelGroupsLds];
  __shared__ double pubv[kNumPub];
  __shared__ double red_v[2 * (NT / 64)];
  __shared__ int red_i[2 * (NT / 64)];

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
  // With xcd_stride 8 the launch is spread over the 8 XCDs (workgroups
  // are dealt round-robin) and only every 8th works, so that all working
  // groups can share one XCD's L2.
  const int stride = p.xcd_stride;
  if (stride > 1 && blockIdx.x % stride != 0) return;
  const int gi = blockIdx.x / stride;  // working-group index
  const int G = gridDim.x / stride;
  const int tid = threadIdx.x;
  const int c0 = s + gi * NT;  // first column owned by this group
  const int j = c0 + tid;
  const bool own = j < n;
  const int bt = min(p.b, n - s);
  T* Wg = reinterpret_cast<T*>(p.Wg);
  // record areas: [slots G][pubs 16][pubcols 3 BM][colR G BM], two parities
  const int PS = G + kNumPub + 3 * BM + G * BM;
  XRec* const xbase = p.xrec;
  int ex = p.epoch;  // exchange number = record tag
  int par = 0;       // reduction scratch half
  int lp = 0;        // Lbuf index of the current pivot

  auto area = [&](int tag) -> XRec* { return xbase + (tag & 1) * PS; };
  auto is_mine = [&](int col) { return col >= c0 && col < c0 + NT && col < n; };

  // L2-scope exchanges only when every working group reports the same
  // XCC_ID (one device-coherent exchange per launch); otherwise every store
  // stays device-coherent.
  bool fast = false;
  if (stride > 1) {
    ++ex;
    XRec* ar = area(ex);
    const int xcc = __builtin_amdgcn_s_getreg((3 << 11) | 20) & 0xf;  // HW_REG_XCC_ID
    if (tid == 0) xput(ar + gi, ex, xcc, 0.0);
    int same = 1;
    for (int id = tid; id < G; id += NT) {
      int aux;
      (void)xwait(ar + id, ex, &aux);
      if (aux != xcc) same = 0;
    }
    fast = __syncthreads_and(same) != 0;
  }
  auto post = [&](XRec* rp, int tag, int aux, double v) {
    if (fast)
      xput_l2(rp, tag, aux, v);
    else
      xput(rp, tag, aux, v);
  };
  auto sta = [&](T* ap, T v) {
    if (fast)
      st_l2<T>(ap, v);
    else
      stc(ap, v);
  };

  T arow = own ? ldc(&A[(int64_t)s * lda + j]) : T(0);
  int kk = 0, nswap = 0;
#if defined(BBK_PANEL_PROFILE)
  // Phase timing with the 100 MHz device clock (build-time option only).
  unsigned long long tp[8] = {0}, pacc[8] = {0};
  int kind_last = 0;
#define BBK_PT(i)                      \
  do {                                 \
    if (tid == 0) tp[i] = wall_clock64(); \
  } while (0)
#else
#define BBK_PT(i) \
  do {            \
  } while (0)
#endif

  while (kk < bt) {
    BBK_PT(0);
    const int k = s + kk;
    const bool has1 = k + 1 < n, has2 = k + 2 < n;
    T* const Lk = Lbuf[lp];
    T* const Lnext = Lbuf[lp ^ 1];
    // ---- phase A: W(k, j) for own j >= k, local argmax over j > k -------
    const bool active = own && j >= k;
    T w = T(0);
    if (active) {
      w = arow - panel_dot<T, NT>(Lk, Wl, tid, kk);
      Wl[kk * NT + tid] = w;
    }
    double v = 0.0;
    int vi = INT_MAX;
    if (active && j > k) {
      const double a = fabs(static_cast<double>(w));
      if (a > 0.0) {  // NaN never wins (strict '>' scan of the reference)
        v = a;
        vi = j;
      }
    }
    // this thread's earlier (coherent) A writes have completed; the barrier
    // of the reduction then orders every thread's before the records below
    wait_vm();
    block_argmax1(v, vi, red_v, red_i, par);
    BBK_PT(1);
    // ---- exchange 1: post --------------------------------------------------
    ++ex;
    {
      XRec* ar = area(ex);
      // slot: local argmax column and its signed value W(k, r_g)
      if (tid == 0)
        post(ar + gi, ex, vi,
             vi != INT_MAX ? static_cast<double>(Wl[kk * NT + (vi - c0)]) : 0.0);
      XRec* pubs = ar + G;
      XRec* pcols = ar + G + kNumPub;
      if (j == k) {
        post(pubs + kPubWk, ex, 0, static_cast<double>(w));
        post(pubs + kPubAkk, ex, 0, static_cast<double>(arow));
      }
      if (j == k + 1) {
        post(pubs + kPubWk1, ex, 0, static_cast<double>(w));
        post(pubs + kPubAkk1, ex, 0, static_cast<double>(arow));
      }
      if (j == k + 2) post(pubs + kPubWk2, ex, 0, static_cast<double>(w));
      if (tid < kk) {
        const int q = tid;
        if (is_mine(k)) post(pcols + q, ex, q, static_cast<double>(Wl[q * NT + (k - c0)]));
        if (has1 && is_mine(k + 1))
          post(pcols + BM + q, ex, q, static_cast<double>(Wl[q * NT + (k + 1 - c0)]));
        if (has2 && is_mine(k + 2))
          post(pcols + 2 * BM + q, ex, q, static_cast<double>(Wl[q * NT + (k + 2 - c0)]));
        if (vi != INT_MAX)
          post(ar + G + kNumPub + 3 * BM + gi * BM + q, ex, q,
               static_cast<double>(Wl[q * NT + (vi - c0)]));
      }
    }
    BBK_PT(2);
    // ---- exchange 1: gather (each thread waits for its own records) --------
    double omega = 0.0;
    int r = INT_MAX;
    {
      const XRec* ar = area(ex);
      const int npub = 2 + (has1 ? 2 : 0) + (has2 ? 1 : 0);
      const int ncol = kk * (1 + (has1 ? 1 : 0) + (has2 ? 1 : 0));
      const int total = G + npub + ncol;
      for (int id = tid; id < total; id += NT) {
        int aux;
        if (id < G) {
          const double val = xwait(ar + id, ex, &aux);
          xval[id] = val;
          argmax_combine(omega, r, fabs(val), aux);
        } else if (id < G + npub) {
          const int u = id - G;
          const int pid = u < 2 ? u : (has1 && u < 4 ? u : kPubWk2);
          pubv[pid] = xwait(ar + G + pid, ex, &aux);
        } else {
          const int u = id - G - npub;
          const int which = u / kk, q = u % kk;  // 0: k, 1: k+1, 2: k+2
          const double val = xwait(ar + G + kNumPub + which * BM + q, ex, &aux);
          if (which == 0)
            colK[q] = val;
          else if (which == 1)
            colK1[q] = val;
          else
            colK2[q] = val;
        }
      }
      block_argmax1(omega, r, red_v, red_i, par);  // also publishes xval/pubv/colK*
    }
    // W(k, r) with its sign, from the slot of r's owner
    const double wkr = (r != INT_MAX) ? xval[(r - s) / NT] : 0.0;
    const double wkk = pubv[kPubWk];
    BBK_PT(3);
    const double akk = fabs(wkk);
    // prefetch rows k+1, k+2 of A for the next pivot (patched after the swap)
    const T P1 = (has1 && own && j >= k + 1) ? ldc(&A[(int64_t)(k + 1) * lda + j]) : T(0);
    const T P2 = (has2 && own && j >= k + 2) ? ldc(&A[(int64_t)(k + 2) * lda + j]) : T(0);
    int kind;
    T c = T(0);
    double crr = 0.0, ck = 0.0, ck1 = 0.0, ck2 = 0.0, a11 = 0.0, a12 = 0.0;
    if (akk >= kAlpha * omega || omega == 0.0) {
      kind = 1;  // src/block_ldlt_decomposition.cc:92
    } else {
      // ---- phase B: candidate row r ----------------------------------------
      const XRec* ar1 = area(ex);
      const int gstar = (r - s) / NT;
      T a = T(0);
      if (active) a = (j < r) ? ldc(&A[(int64_t)j * lda + r]) : ldc(&A[(int64_t)r * lda + j]);
      if (tid < kk) {
        // W(q, r) from r's owner; L(r, s+q) needs the 2x2 partner's value too
        const XRec* cr = ar1 + G + kNumPub + 3 * BM + gstar * BM;
        int aux;
        const double cq = xwait(cr + tid, ex, &aux);
        const int ty = dt[tid];
        double partner = 0.0;
        if (ty == 2)
          partner = xwait(cr + tid + 1, ex, &aux);
        else if (ty == -2)
          partner = xwait(cr + tid - 1, ex, &aux);
        colR[tid] = cq;
        Lr[tid] = static_cast<T>(ucol_pair(ty, dc0[tid], dc1[tid], cq, partner));
      }
      __syncthreads();
      if (active) {
        c = a - panel_dot<T, NT>(Lr, Wl, tid, kk);
      }
      double sv = 0.0;
      if (active && j != r) {
        const double aa = fabs(static_cast<double>(c));
        sv = (aa >= 0.0) ? aa : 0.0;  // NaN ignored, as std::max does
      }
      sv = block_max1(sv, red_v, par);
      BBK_PT(4);
      // ---- exchange 2 ------------------------------------------------------
      // Rows k+1/k+2 prefetches may still be in flight: the elements other
      // groups can overwrite in phase C are exactly the ones phase C
      // replaces by published values, so no wait is needed here.
      ++ex;
      XRec* ar = area(ex);
      XRec* pubs = ar + G;
      if (tid == 0) post(ar + gi, ex, 0, sv);
      if (j == r) post(pubs + kPubCr, ex, 0, static_cast<double>(c));
      if (j == k) post(pubs + kPubCk, ex, 0, static_cast<double>(c));
      if (j == k + 1) {
        post(pubs + kPubCk1, ex, 0, static_cast<double>(c));
        post(pubs + kPubA11, ex, 0, static_cast<double>(P1));
      }
      if (j == k + 2) {
        post(pubs + kPubCk2, ex, 0, static_cast<double>(c));
        post(pubs + kPubA12, ex, 0, static_cast<double>(P1));
      }
      const int npub = 2 + (has1 ? 2 : 0) + (has2 ? 2 : 0);
      const int total = G + npub;
      double sg = 0.0;
      for (int id = tid; id < total; id += NT) {
        int aux;
        if (id < G) {
          const double o = xwait(ar + id, ex, &aux);
          sg = (sg < o) ? o : sg;
        } else {
          const int pid = kPubCr + (id - G);  // 8, 9, (10, 11), (12, 13)
          pubv[pid] = xwait(pubs + pid, ex, &aux);
        }
      }
      const double sigma = block_max1(sg, red_v, par);  // also publishes pubv
      crr = pubv[kPubCr];
      ck = pubv[kPubCk];
      if (has1) {
        ck1 = pubv[kPubCk1];
        a11 = pubv[kPubA11];
      }
      if (has2) {
        ck2 = pubv[kPubCk2];
        a12 = pubv[kPubA12];
      }
      kind = select_pivot_nc(akk, omega, sigma, fabs(crr));
      BBK_PT(5);
    }
#if defined(BBK_PANEL_PROFILE)
    kind_last = kind;
#endif

    // ---- phase C: interchange, scale, and set up the next pivot ------------
    int size;
    T arow_next = P1;
    if (kind == 1 || kind == 3) {
      // ScaleOneByOnePivot, src/block_ldlt_decomposition.cc:633-656
      const double den = safe_den(wkk);
      if (own && j > k)
        sta(&A[(int64_t)k * lda + j], static_cast<T>(div_or_zero(static_cast<double>(w), den)));
      if (j == k) {
        sta(&A[(int64_t)k * lda + k], w);
        p.piv[k] = 1;
      }
      if (tid == 0) {
        dt[kk] = 1;
        dc0[kk] = den;
        dc1[kk] = 0.0;
      }
      if (has1) {
        if (tid < kk) Lnext[tid] = static_cast<T>(ucol(dt, dc0, dc1, colK1, tid));
        if (tid == kk) Lnext[kk] = static_cast<T>(div_or_zero(pubv[kPubWk1], den));
      }
      size = 1;
    } else if (kind == -1) {
      // symmetric interchange k <-> r (src/block_ldlt_decomposition.cc:414-498)
      // columns k and r of the finished panel rows, one row per thread
      if (tid < kk) {
        const int q = tid;
        if (is_mine(k)) {
          Wl[q * NT + (k - c0)] = static_cast<T>(colR[q]);
          sta(&A[(int64_t)(s + q) * lda + k], Lr[q]);
        }
        if (is_mine(r)) {
          Wl[q * NT + (r - c0)] = static_cast<T>(colK[q]);
          sta(&A[(int64_t)(s + q) * lda + r], Lk[q]);
        }
      }
      T nw = c;
      if (j == k)
        nw = static_cast<T>(crr);
      else if (j == r)
        nw = static_cast<T>(ck);
      if (active) Wl[kk * NT + tid] = nw;
      if (own && j > r)
        sta(&A[(int64_t)r * lda + j], arow);
      else if (own && j > k && j < r)
        sta(&A[(int64_t)j * lda + r], arow);
      if (j == r) sta(&A[(int64_t)r * lda + r], static_cast<T>(pubv[kPubAkk]));
      const double den = safe_den(crr);
      if (own && j > k)
        sta(&A[(int64_t)k * lda + j], static_cast<T>(div_or_zero(static_cast<double>(nw), den)));
      if (j == k) {
        sta(&A[(int64_t)k * lda + k], nw);
        p.piv[k] = 1;
      }
      if (gi == 0 && tid == 0) {
        const int t = p.perm[k];
        p.perm[k] = p.perm[r];
        p.perm[r] = t;
        p.swaps[2 * nswap] = k;
        p.swaps[2 * nswap + 1] = r;
      }
      ++nswap;
      if (tid == 0) {
        dt[kk] = 1;
        dc0[kk] = den;
        dc1[kk] = 0.0;
      }
      if (has1) {
        if (r == k + 1) {
          if (tid < kk) Lnext[tid] = Lk[tid];
          if (tid == kk) Lnext[kk] = static_cast<T>(div_or_zero(ck, den));
          arow_next = (j == k + 1) ? static_cast<T>(pubv[kPubAkk]) : arow;
        } else {
          if (tid < kk) Lnext[tid] = static_cast<T>(ucol(dt, dc0, dc1, colK1, tid));
          if (tid == kk) Lnext[kk] = static_cast<T>(div_or_zero(ck1, den));
          if (j == r) arow_next = static_cast<T>(pubv[kPubAkk1]);
        }
      }
      size = 1;
    } else {
      // 2x2 pivot at (k, k+1) after the interchange k+1 <-> r
      // (src/block_ldlt_decomposition.cc:500-587, 658-704)
      const int k1 = k + 1;
      if (r != k1) {
        // columns k+1 and r of the finished panel rows, one row per thread
        if (tid < kk) {
          const int q = tid;
          if (is_mine(k1)) {
            Wl[q * NT + (k1 - c0)] = static_cast<T>(colR[q]);
            sta(&A[(int64_t)(s + q) * lda + k1], Lr[q]);
          }
          if (is_mine(r)) {
            Wl[q * NT + (r - c0)] = static_cast<T>(colK1[q]);
            sta(&A[(int64_t)(s + q) * lda + r], static_cast<T>(ucol(dt, dc0, dc1, colK1, q)));
          }
        }
        // trailing matrix: new row/column r take the old row k+1 (= P1)
        if (own && j > r)
          sta(&A[(int64_t)r * lda + j], P1);
        else if (own && j > k1 && j < r)
          sta(&A[(int64_t)j * lda + r], P1);
        if (j == r) sta(&A[(int64_t)r * lda + r], static_cast<T>(a11));
        if (gi == 0 && tid == 0) {
          const int t = p.perm[k1];
          p.perm[k1] = p.perm[r];
          p.perm[r] = t;
          p.swaps[2 * nswap] = k1;
          p.swaps[2 * nswap + 1] = r;
        }
        ++nswap;
      }
      // new W rows kk (current row) and kk+1 (candidate row), columns
      // k+1 and r exchanged
      T w0 = active ? Wl[kk * NT + tid] : T(0);
      T w1 = c;
      if (j == k1) {
        w0 = static_cast<T>(wkr);
        w1 = static_cast<T>(crr);
      } else if (j == r) {
        w0 = static_cast<T>(pubv[kPubWk1]);
        w1 = static_cast<T>(ck1);
      }
      if (active) {
        Wl[kk * NT + tid] = w0;
        Wl[(kk + 1) * NT + tid] = w1;
      }
      const double a = wkk, bb = wkr, cc = crr;
      double itl, ibr, ioff;
      inverse_2x2_nc(a, bb, cc, &itl, &ibr, &ioff);
      if (own && j >= k + 2) {
        double u0, u1;
        apply_2x2_nc(itl, ibr, ioff, static_cast<double>(w0), static_cast<double>(w1), &u0,
                     &u1);
        sta(&A[(int64_t)k * lda + j], static_cast<T>(u0));
        sta(&A[(int64_t)k1 * lda + j], static_cast<T>(u1));
      }
      if (j == k) {
        sta(&A[(int64_t)k * lda + k], static_cast<T>(a));
        p.piv[k] = 2;
      }
      if (j == k1) {
        sta(&A[(int64_t)k * lda + k1], static_cast<T>(bb));
        sta(&A[(int64_t)k1 * lda + k1], static_cast<T>(cc));
        p.piv[k1] = -2;
      }
      if (tid == 0) {
        dt[kk] = 2;
        dc0[kk] = itl;
        dc1[kk] = ioff;
        dt[kk + 1] = -2;
        dc0[kk + 1] = ioff;
        dc1[kk + 1] = ibr;
      }
      if (has2) {
        const bool rk2 = (r == k + 2);
        const double* src = rk2 ? colK1 : colK2;
        if (tid < kk) Lnext[tid] = static_cast<T>(ucol(dt, dc0, dc1, src, tid));
        const double wa = rk2 ? pubv[kPubWk1] : pubv[kPubWk2];
        const double wb = rk2 ? ck1 : ck2;
        double u0, u1;
        apply_2x2_nc(itl, ibr, ioff, wa, wb, &u0, &u1);
        if (tid == kk) Lnext[kk] = static_cast<T>(u0);
        if (tid == kk + 1) Lnext[kk + 1] = static_cast<T>(u1);
        if (rk2)
          arow_next = (j == k + 2) ? static_cast<T>(a11) : P1;
        else
          arow_next = (j == r && r > k + 2) ? static_cast<T>(a12) : P2;
      }
      size = 2;
    }
    __syncthreads();  // Lnext, dt/dc, Wl of this pivot complete
    lp ^= 1;
    arow = arow_next;
    kk += size;
#if defined(BBK_PANEL_PROFILE)
    BBK_PT(6);
    if (tid == 0) {
      pacc[0] += tp[1] - tp[0];
      pacc[1] += tp[2] - tp[1];
      pacc[2] += tp[3] - tp[2];
      if (kind_last != 1) {
        pacc[3] += tp[4] - tp[3];
        pacc[4] += tp[5] - tp[4];
        pacc[5] += tp[6] - tp[5];
        pacc[7] += 1;
      } else {
        pacc[5] += tp[6] - tp[3];
        pacc[6] += 1;
      }
    }
#endif
  }
#if defined(BBK_PANEL_PROFILE)
  if (tid == 0 && (gi == 0 || gi == G - 1))
    for (int i = 0; i < 8; ++i) atomicAdd(&g_panel_profile[(gi == 0 ? 0 : 8) + i], pacc[i]);
#endif

  if (own)
    for (int q = 0; q < kk; ++q) Wg[(int64_t)q * p.ldw + j] = Wl[q * NT + tid];
  if (gi == 0 && tid == 0) {
    p.state[PanelState::kStart] = s;
    p.state[PanelState::kWidth] = kk;
    p.state[PanelState::kNumSwaps] = nswap;
    p.state[PanelState::kNextStart] = s + kk;
  }
}

}  // namespace bbk

namespace bbk {

template <typename T>
hipError_t launch_panel(T* A, const PanelArgs& args, int groups, hipStream_t stream) {
  T* a = A;
  PanelArgs pa = args;
  void* kargs[] = {&a, &pa};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&bk_panel_kernel<T>),
                                    dim3(groups), dim3(kPanelThreads), kargs, 0, stream);
}

template <typename T>
hipError_t panel_max_groups(int* groups) {
  int per_cu = 0, dev = 0, cus = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(&bk_panel_kernel<T>), kPanelThreads, 0);
  if (e != hipSuccess) return e;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  *groups = per_cu * cus;
  return hipSuccess;
}

size_t panel_exchange_records(int groups, int elem_bytes) {
  const int bm = elem_bytes == 8 ? PanelCfg<double>::kMaxWidth : PanelCfg<float>::kMaxWidth;
  return 2 * static_cast<size_t>(groups + 16 + 3 * bm + groups * bm);
}

template hipError_t launch_panel<double>(double*, const PanelArgs&, int, hipStream_t);
template hipError_t launch_panel<float>(float*, const PanelArgs&, int, hipStream_t);
template hipError_t panel_max_groups<double>(int*);
template hipError_t panel_max_groups<float>(int*);

}  // namespace bbk

#if defined(BBK_PANEL_PROFILE)
// Read (and optionally clear) the phase clock totals of the profiling build.
extern "C" int bbk_panel_profile(unsigned long long* out, int reset) {
  hipError_t e = hipMemcpyFromSymbol(out, HIP_SYMBOL(bbk::g_panel_profile),
                                     sizeof(unsigned long long) * 16);
  if (reset) {
    unsigned long long z[16] = {0};
    (void)hipMemcpyToSymbol(HIP_SYMBOL(bbk::g_panel_profile), z, sizeof(z));
  }
  return e == hipSuccess ? 0 : 1;
}
#endif

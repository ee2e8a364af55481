// bbk_panel.hip -- the Bunch-Kaufman panel kernel for gfx950.
//
// One cooperative launch factors one panel of up to b columns starting at
// s (read from device state).  Workgroups of NT threads (256, or 512 for
// orders past 65536) hold NT active columns each, in 64-column chunks dealt
// round-robin over the groups (below); thread t holds one column
// (position) and keeps that column of the panel's unscaled rows W = D L^T
// in LDS (the first kRegRows rows -- 64, or 32 in the 512-thread binary64
// variant -- move to registers once the panel passes them).  An interchange swaps the positions of the two threads
// holding the interchanged columns; no W entry ever moves.
//
// Per pivot k (local index kk) the workgroups meet at most twice:
//   exchange 1  after W(k, :) = A(k, :) - L(k, panel) W(panel, :) and the
//               local argmax; carries (omega_g, r_g, W(k, r_g)) and a few
//               scalars;
//   exchange 2  (only when the Bunch-Kaufman test needs the candidate row r)
//               after the candidate row and its local max sigma_g; carries
//               sigma_g and the candidate values at k, k+1, k+2, r.
// Every datum travels as a 16-byte record {tag, aux, value} written and
// read with device-coherent (sc1) vector accesses; a reader accepts a
// record only when its tag equals the exchange's number, so an exchange is
// one store per writer and one polling read per reader, with no atomics.
// Two record areas alternate by exchange parity.  Everything the next
// pivot needs (its A row, its L row) is rebuilt locally from published
// values and from A, so no third synchronization per pivot is needed: the
// L rows of the candidate and of the next pivot are columns of the finished
// U rows, read back from A.
//
// Arithmetic: pivot tests, 2x2 inverse and L = D^-1 W in binary64 without
// contraction (src/block_ldlt_decomposition.cc:21-26,369-390,633-704).
#include <climits>

#include "bbk_kernels.hpp"

namespace bbk {

#if defined(BBK_PANEL_PROFILE)
// Per-phase clock totals over the pivots that need the candidate row, of
// groups 0 (slots 0-11) and G-1 (slots 12-23): slots 0-8 the intervals
// between the BBK_PT points 0..9, slot 9 the number of such pivots.
__device__ unsigned long long g_panel_profile[24];
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
__device__ __forceinline__ void stc(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Device-coherent load the compiler sees and waits for.  Used in phase B:
// there every value loaded is consumed before the next store of the pivot,
// so the compiler's waits never hold back a store, and it can wait for the
// older loads (vmcnt(n) > 0) while the younger ones are still in flight.
template <typename T>
__device__ __forceinline__ T ldc(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Device-coherent loads issued together, with the wait for them in the same
// asm block, for loads whose values live across the pivot's stores (the
// next pivot's row and L row on the 1x1 path, the first row): a load the
// compiler sees in flight makes it wait, where control flow merges, for
// every store issued since (one counter on gfx9 covers loads and stores),
// which would serialize the pivot's coherent stores.  Every pointer must be
// valid; lanes that need no value pass a harmless address.
template <typename T>
struct CoherentLoad;
template <>
struct CoherentLoad<double> {
  static __device__ __forceinline__ void one(const double* p0, double& v0) {
    asm volatile("global_load_dwordx2 %0, %1, off sc1\n\ts_waitcnt vmcnt(0)"
                 : "=&v"(v0)
                 : "v"(p0)
                 : "memory");
  }
  static __device__ __forceinline__ void two(const double* p0, const double* p1, double& v0,
                                             double& v1) {
    asm volatile(
        "global_load_dwordx2 %0, %2, off sc1\n\t"
        "global_load_dwordx2 %1, %3, off sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(v0), "=&v"(v1)
        : "v"(p0), "v"(p1)
        : "memory");
  }
};
template <>
struct CoherentLoad<float> {
  static __device__ __forceinline__ void one(const float* p0, float& v0) {
    asm volatile("global_load_dword %0, %1, off sc1\n\ts_waitcnt vmcnt(0)"
                 : "=&v"(v0)
                 : "v"(p0)
                 : "memory");
  }
  static __device__ __forceinline__ void two(const float* p0, const float* p1, float& v0,
                                             float& v1) {
    asm volatile(
        "global_load_dword %0, %2, off sc1\n\t"
        "global_load_dword %1, %3, off sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(v0), "=&v"(v1)
        : "v"(p0), "v"(p1)
        : "memory");
  }
};

// sum_q L[q] * W(q, own column), q < kk.  Rows [0, RR) come from the
// registers wr once base == RR; rows [base, kk) from LDS, row q at
// (q - base) * WS.  Four accumulators keep several loads in flight.
template <typename T, int WS, int RR, int NR>
__device__ __forceinline__ T panel_dot(const T* L, const T* Wl, const T (&wr)[NR], int tid,
                                       int kk, int base) {
  T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
  if constexpr (RR > 0) {
    if (base > 0) {
      // L in blocks of 16 read as 16-byte vectors, all loads of a block
      // issued before its products
      static_assert(RR % 16 == 0, "register rows come in blocks of 16");
      constexpr int VE = 16 / sizeof(T);  // elements per vector
      typedef T vec_t __attribute__((ext_vector_type(VE)));
      const vec_t* Lv = reinterpret_cast<const vec_t*>(L);
#pragma unroll
      for (int b = 0; b < RR; b += 16) {
        T lv[16];
#pragma unroll
        for (int u = 0; u < 16 / VE; ++u) {
          const vec_t t = Lv[b / VE + u];
#pragma unroll
          for (int e = 0; e < VE; ++e) lv[u * VE + e] = t[e];
        }
#pragma unroll
        for (int u = 0; u < 16; u += 4) {
          a0 += lv[u] * wr[b + u];
          a1 += lv[u + 1] * wr[b + u + 1];
          a2 += lv[u + 2] * wr[b + u + 2];
          a3 += lv[u + 3] * wr[b + u + 3];
        }
      }
    }
  }
  const T* Wc = Wl + tid;
  const T* Lb = L + base;  // 16-byte aligned: L is, and base is 0 or 64
  const int cnt = kk - base;
  int q = 0;
  // blocks of 8 rows: all 16 loads of a block are issued before its
  // products (one LDS round trip per block)
  for (; q + 8 <= cnt; q += 8) {
    T lv[8], wv[8];
    if constexpr (sizeof(T) == 8) {
      typedef double d2 __attribute__((ext_vector_type(2)));
      const d2* L2 = reinterpret_cast<const d2*>(Lb + q);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const d2 t = L2[u];
        lv[2 * u] = t.x;
        lv[2 * u + 1] = t.y;
      }
    } else {
      typedef float f4 __attribute__((ext_vector_type(4)));
      const f4* L4 = reinterpret_cast<const f4*>(Lb + q);
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const f4 t = L4[u];
        lv[4 * u] = t.x;
        lv[4 * u + 1] = t.y;
        lv[4 * u + 2] = t.z;
        lv[4 * u + 3] = t.w;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) wv[u] = Wc[(q + u) * WS];
    a0 += lv[0] * wv[0];
    a1 += lv[1] * wv[1];
    a2 += lv[2] * wv[2];
    a3 += lv[3] * wv[3];
    a0 += lv[4] * wv[4];
    a1 += lv[5] * wv[5];
    a2 += lv[6] * wv[6];
    a3 += lv[7] * wv[7];
  }
  for (; q < cnt; ++q) a0 += Lb[q] * Wc[q * WS];
  return (a0 + a1) + (a2 + a3);
}

// Three products at once over q < kk (W read once): d0, d1, d2 with L0,
// L1, L2.  Binary32 forms the candidate row's product and both possible next
// rows' products together in phase B, under the candidate element's HBM
// load, instead of the next rows' during exchange 2.
template <typename T, int WS, int RR, int NR>
__device__ __forceinline__ void panel_dot3(const T* L0, const T* L1, const T* L2, const T* Wl,
                                           const T (&wr)[NR], int tid, int kk, int base, T& d0,
                                           T& d1, T& d2) {
  T a0 = T(0), a1 = T(0), b0 = T(0), b1 = T(0), c0 = T(0), c1 = T(0);
  constexpr int VE = 16 / sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(VE)));
  if constexpr (RR > 0) {
    if (base > 0) {
      const vec_t* V0 = reinterpret_cast<const vec_t*>(L0);
      const vec_t* V1 = reinterpret_cast<const vec_t*>(L1);
      const vec_t* V2 = reinterpret_cast<const vec_t*>(L2);
#pragma unroll
      for (int b = 0; b < RR; b += 8) {
        T l0[8], l1[8], l2[8];
#pragma unroll
        for (int u = 0; u < 8 / VE; ++u) {
          const vec_t t0 = V0[b / VE + u], t1 = V1[b / VE + u], t2 = V2[b / VE + u];
#pragma unroll
          for (int e = 0; e < VE; ++e) {
            l0[u * VE + e] = t0[e];
            l1[u * VE + e] = t1[e];
            l2[u * VE + e] = t2[e];
          }
        }
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
          a0 += l0[u] * wr[b + u];
          a1 += l0[u + 1] * wr[b + u + 1];
          b0 += l1[u] * wr[b + u];
          b1 += l1[u + 1] * wr[b + u + 1];
          c0 += l2[u] * wr[b + u];
          c1 += l2[u + 1] * wr[b + u + 1];
        }
      }
    }
  }
  const T* Wc = Wl + tid;
  const T* L0b = L0 + base;
  const T* L1b = L1 + base;
  const T* L2b = L2 + base;
  const int cnt = kk - base;
  int q = 0;
  for (; q + 8 <= cnt; q += 8) {
    T l0[8], l1[8], l2[8], wv[8];
    const vec_t* V0 = reinterpret_cast<const vec_t*>(L0b + q);
    const vec_t* V1 = reinterpret_cast<const vec_t*>(L1b + q);
    const vec_t* V2 = reinterpret_cast<const vec_t*>(L2b + q);
#pragma unroll
    for (int u = 0; u < 8 / VE; ++u) {
      const vec_t t0 = V0[u], t1 = V1[u], t2 = V2[u];
#pragma unroll
      for (int e = 0; e < VE; ++e) {
        l0[u * VE + e] = t0[e];
        l1[u * VE + e] = t1[e];
        l2[u * VE + e] = t2[e];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) wv[u] = Wc[(q + u) * WS];
#pragma unroll
    for (int u = 0; u < 8; u += 2) {
      a0 += l0[u] * wv[u];
      a1 += l0[u + 1] * wv[u + 1];
      b0 += l1[u] * wv[u];
      b1 += l1[u + 1] * wv[u + 1];
      c0 += l2[u] * wv[u];
      c1 += l2[u + 1] * wv[u + 1];
    }
  }
  for (; q < cnt; ++q) {
    a0 += L0b[q] * Wc[q * WS];
    b0 += L1b[q] * Wc[q * WS];
    c0 += L2b[q] * Wc[q * WS];
  }
  d0 = a0 + a1;
  d1 = b0 + b1;
  d2 = c0 + c1;
}

// Two products at once over q < kk: d1 = sum_q L1[q] W(q, own column),
// d2 = sum_q L2[q] W(q, own column) (same storage as panel_dot; W is read
// once).  Used for the next pivot's row while an exchange is in flight.
template <typename T, int WS, int RR, int NR>
__device__ __forceinline__ void panel_dot2(const T* L1, const T* L2, const T* Wl,
                                           const T (&wr)[NR], int tid, int kk, int base, T& d1,
                                           T& d2) {
  T a0 = T(0), a1 = T(0), b0 = T(0), b1 = T(0);
  constexpr int VE = 16 / sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(VE)));
  if constexpr (RR > 0) {
    if (base > 0) {
      const vec_t* V1 = reinterpret_cast<const vec_t*>(L1);
      const vec_t* V2 = reinterpret_cast<const vec_t*>(L2);
#pragma unroll
      for (int b = 0; b < RR; b += 8) {
        T l1[8], l2[8];
#pragma unroll
        for (int u = 0; u < 8 / VE; ++u) {
          const vec_t t1 = V1[b / VE + u], t2 = V2[b / VE + u];
#pragma unroll
          for (int e = 0; e < VE; ++e) {
            l1[u * VE + e] = t1[e];
            l2[u * VE + e] = t2[e];
          }
        }
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
          a0 += l1[u] * wr[b + u];
          a1 += l1[u + 1] * wr[b + u + 1];
          b0 += l2[u] * wr[b + u];
          b1 += l2[u + 1] * wr[b + u + 1];
        }
      }
    }
  }
  const T* Wc = Wl + tid;
  const T* L1b = L1 + base;  // 16-byte aligned: L1, L2 are, base is 0 or RR
  const T* L2b = L2 + base;
  const int cnt = kk - base;
  int q = 0;
  for (; q + 8 <= cnt; q += 8) {
    T l1[8], l2[8], wv[8];
    const vec_t* V1 = reinterpret_cast<const vec_t*>(L1b + q);
    const vec_t* V2 = reinterpret_cast<const vec_t*>(L2b + q);
#pragma unroll
    for (int u = 0; u < 8 / VE; ++u) {
      const vec_t t1 = V1[u], t2 = V2[u];
#pragma unroll
      for (int e = 0; e < VE; ++e) {
        l1[u * VE + e] = t1[e];
        l2[u * VE + e] = t2[e];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) wv[u] = Wc[(q + u) * WS];
#pragma unroll
    for (int u = 0; u < 8; u += 2) {
      a0 += l1[u] * wv[u];
      a1 += l1[u + 1] * wv[u + 1];
      b0 += l2[u] * wv[u];
      b1 += l2[u + 1] * wv[u + 1];
    }
  }
  for (; q < cnt; ++q) {
    a0 += L1b[q] * Wc[q * WS];
    b0 += L2b[q] * Wc[q * WS];
  }
  d1 = a0 + a1;
  d2 = b0 + b1;
}

enum PubId {
  kPubWk = 0, kPubAkk = 1, kPubWk1 = 2, kPubAkk1 = 3, kPubWk2 = 4, kPubWkr = 5,
  kPubCr = 8, kPubCk = 9, kPubCk1 = 10, kPubA11 = 11, kPubCk2 = 12, kPubA12 = 13,
  kNumPub = 16
};

template <typename T, int NT>
__global__ __launch_bounds__(NT) void bk_panel_kernel(T* __restrict__ A, PanelArgs p) {
  constexpr int BM = PanelCfg<T, NT>::kMaxWidth;
  constexpr int RR = PanelCfg<T, NT>::kRegRows;
  constexpr int LR = PanelCfg<T, NT>::kLdsRows;
  static_assert(NT % 64 == 0 && NT >= BM + 2, "a thread per L-row entry of the panel");
  constexpr int NR = RR > 0 ? RR : 1;
  // rows [0, RR], RR + 1 of them, sit in Wl until they move to registers
  static_assert(LR >= RR, "the LDS rows must hold the register rows before they move");
  constexpr int WS = NT;  // LDS row stride of W
  // W(q, own column): q < base (= 0 or RR) in registers wr[q], the rest in
  // LDS row q - base
  __shared__ T Wl[(LR + 1) * WS];
  // L(k, panel) of this pivot / of the next, and of the candidate row
  __shared__ __attribute__((aligned(16))) T Lbuf[2][BM + 2];
  __shared__ __attribute__((aligned(16))) T Lr[BM + 2];
  // U(q, k+1), U(q, k+2), q < kk: the candidates for the next pivot's L row
  __shared__ __attribute__((aligned(16))) T Lu1[BM + 2];
  __shared__ __attribute__((aligned(16))) T Lu2[BM + 2];
  __shared__ double pubv[kNumPub];
  __shared__ double red_v[2 * (NT / 64)];
  __shared__ int red_i[2 * (NT / 64)];
  // this panel's interchanges (workgroup 0), applied to perm at the end
  __shared__ int swl[2 * (BM + 1)], sw_slot[2 * (BM + 1)], sw_content[2 * (BM + 1)];

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
  // Lane l of wavefront w starts at position (column) j = s + 64 (w G + gi)
  // + l: the 64-column chunks are dealt to the groups round-robin, so every
  // group holds columns across the whole active range and the candidate
  // row's strided part (the columns left of r) is spread evenly.  A thread
  // keeps its W column for the whole panel: a symmetric interchange of
  // positions a and b swaps the positions of the two threads holding them
  // instead of moving W entries.
  int j = s + 64 * ((tid >> 6) * G + gi) + (tid & 63);
  const bool own = j < n;
  const int qrow = gi + tid * G;  // finished panel row whose interchanges this thread writes
  const int bt = min(p.b, n - s);
  T* Wg = reinterpret_cast<T*>(p.Wg);
  // record areas: [slots G][pubs 16], two parities
  const int PS = G + kNumPub;
  XRec* const xbase = p.xrec;
  int ex = p.epoch;  // exchange number = record tag
  int par = 0;       // reduction scratch half
  int lp = 0;        // Lbuf index of the current pivot

  auto area = [&](int tag) -> XRec* { return xbase + (tag & 1) * PS; };
  const WaitGuard wg(p.status, p.wait_ticks);

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
      (void)xwait(ar + id, ex, &aux, wg);
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

  // a harmless address for lanes that load nothing (a valid element)
  const T* const dummy = A + (int64_t)s * lda + s;
  T arow;
  CoherentLoad<T>::one(own ? &A[(int64_t)s * lda + j] : dummy, arow);
  if (!own) arow = T(0);
  int kk = 0, nswap = 0;
  // the next pivot's row product, when it was formed during exchange 2
  bool have_next = false;
  T pnext = T(0);
  T wr[NR];
  int base = 0;
  // interchange of positions a < b: their holders swap positions
  auto swap_pos = [&](int a, int b) {
    if (j == a)
      j = b;
    else if (j == b)
      j = a;
  };
#if defined(BBK_PANEL_PROFILE)
  // Phase timing with the 100 MHz device clock (build-time option only).
  unsigned long long tp[10] = {0}, pacc[12] = {0};
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
    // a pivot at k+2 can follow in this panel
    const bool use2 = has2 && kk + 2 < bt;
    if constexpr (RR > 0) {
      if (base == 0 && kk >= RR) {
        // rows [0, RR) to registers; a row RR (2x2 across the boundary)
        // moves to LDS row 0
#pragma unroll
        for (int q = 0; q < RR; ++q) wr[q] = Wl[q * WS + tid];
        for (int q = RR; q < kk; ++q) Wl[(q - RR) * WS + tid] = Wl[q * WS + tid];
        base = RR;
        __syncthreads();
      }
    }
    T* const Lk = Lbuf[lp];
    T* const Lnext = Lbuf[lp ^ 1];
    // ---- phase A: W(k, j) for own j >= k, local argmax over j > k -------
    const bool active = own && j >= k;
    T w = T(0), dotA = T(0);
    if (active) {
      dotA = have_next ? pnext : panel_dot<T, WS, RR>(Lk, Wl, wr, tid, kk, base);
      w = arow - dotA;
      Wl[(kk - base) * WS + tid] = w;
    }
    have_next = false;
    BBK_PT(1);
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
    BBK_PT(2);
    block_argmax1(v, vi, red_v, red_i, par);
    BBK_PT(3);
    // ---- exchange 1: post --------------------------------------------------
    ++ex;
    {
      XRec* ar = area(ex);
      // slot: local argmax position and its signed value W(k, r_g), posted
      // by its holder
      if (vi != INT_MAX ? j == vi : tid == 0)
        post(ar + gi, ex, vi, vi != INT_MAX ? static_cast<double>(w) : 0.0);
      XRec* pubs = ar + G;
      if (j == k) {
        post(pubs + kPubWk, ex, 0, static_cast<double>(w));
        post(pubs + kPubAkk, ex, 0, static_cast<double>(arow));
      }
      if (j == k + 1) {
        post(pubs + kPubWk1, ex, 0, static_cast<double>(w));
        post(pubs + kPubAkk1, ex, 0, static_cast<double>(arow));
      }
      if (j == k + 2) post(pubs + kPubWk2, ex, 0, static_cast<double>(w));
    }
    BBK_PT(4);
    // ---- exchange 1: gather (each thread waits for its own records) --------
    double omega = 0.0;
    int r = INT_MAX;
    {
      const XRec* ar = area(ex);
      const int npub = 2 + (has1 ? 2 : 0) + (has2 ? 1 : 0);
      const int total = G + npub;
      int slot_aux = INT_MAX;
      double slot_val = 0.0;
      for (int id = tid; id < total; id += NT) {
        int aux;
        if (id < G) {
          slot_val = xwait(ar + id, ex, &aux, wg);
          slot_aux = aux;
          argmax_combine(omega, r, fabs(slot_val), aux);
        } else {
          const int u = id - G;
          const int pid = u < 2 ? u : (has1 && u < 4 ? u : kPubWk2);
          pubv[pid] = xwait(ar + G + pid, ex, &aux, wg);
        }
      }
      block_argmax1(omega, r, red_v, red_i, par);  // also publishes pubv
      // W(k, r) with its sign, from the slot naming r (read after a later
      // barrier: only the 2x2 pivot uses it)
      if (r != INT_MAX && slot_aux == r) pubv[kPubWkr] = slot_val;
    }
    const double wkk = pubv[kPubWk];
    BBK_PT(5);
    const double akk = fabs(wkk);
    // prefetch row k+1 of A for the next pivot (patched after the swap)
    const bool want1 = has1 && own && j >= k + 1;
    const bool want2 = has2 && own && j >= k + 2;  // row k+2: only after a 2x2 pivot
    const T* const p1 = want1 ? &A[(int64_t)(k + 1) * lda + j] : dummy;
    // L(k+1, s+q) = U(q, k+1) (and U(q, k+2)), q = tid < kk: the next
    // pivot's L row from the finished rows as stored (bit for bit the
    // values D^-1 W(q, .) the scaling steps wrote, also after interchanges,
    // which move stored columns and W columns together)
    const bool wantu = tid < kk;
    const T* const pu1 = (wantu && has1) ? &A[(int64_t)(s + tid) * lda + k + 1] : dummy;
    T P1, P2 = T(0), u1 = T(0), u2 = T(0);
    int kind;
    T c = T(0), pd1 = T(0), pd2 = T(0);
    double crr = 0.0, ck = 0.0, ck1 = 0.0, ck2 = 0.0, a11 = 0.0, a12 = 0.0;
    if (akk >= kAlpha * omega || omega == 0.0) {
      kind = 1;  // src/block_ldlt_decomposition.cc:92
      CoherentLoad<T>::two(p1, pu1, P1, u1);
      if (!want1) P1 = T(0);
    } else {
      // ---- phase B: candidate row r ----------------------------------------
      // candidate row element, L(r, s+q) = U(q, r) for q = tid < kk (column r
      // of the finished rows as stored), and for the next pivot the rows
      // k+1, k+2 and the L rows U(q, k+1), U(q, k+2)
      const T* pa = active ? ((j < r) ? &A[(int64_t)j * lda + r] : &A[(int64_t)r * lda + j])
                           : dummy;
      const T* pl = wantu ? &A[(int64_t)(s + tid) * lda + r] : dummy;
      const T* p2 = want2 ? &A[(int64_t)(k + 2) * lda + j] : dummy;
      const T* pu2 = (wantu && use2) ? &A[(int64_t)(s + tid) * lda + k + 2] : dummy;
      // The finished rows' columns (recently written, cache-resident) are
      // issued first and the candidate element and rows k+1, k+2 (HBM) last:
      // loads return in issue order, so the product over the finished rows
      // runs while the HBM loads are still in flight.
      const T lr = ldc(pl);
      u1 = ldc(pu1);
      u2 = ldc(pu2);
      T a = ldc(pa);
      P1 = ldc(p1);
      P2 = ldc(p2);
      if (tid < kk) {
        Lr[tid] = lr;
        Lu1[tid] = u1;
        Lu2[tid] = u2;
      }
      __syncthreads();
      BBK_PT(6);
      T dotB = T(0);
      if (active) {
        if constexpr (sizeof(T) == 4)
          panel_dot3<T, WS, RR>(Lr, Lu1, Lu2, Wl, wr, tid, kk, base, dotB, pd1, pd2);
        else
          dotB = panel_dot<T, WS, RR>(Lr, Wl, wr, tid, kk, base);
      }
      if (!active) a = T(0);
      if (!want1) P1 = T(0);
      if (!want2) P2 = T(0);
      if (active) c = a - dotB;
      double sv = 0.0;
      if (active && j != r) {
        const double aa = fabs(static_cast<double>(c));
        sv = (aa >= 0.0) ? aa : 0.0;  // NaN ignored, as std::max does
      }
      sv = block_max1(sv, red_v, par);
      BBK_PT(7);
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
      // while the records travel: the next pivot's row product over the
      // finished rows q < kk for both possible next rows (k+1, k+2)
      // (binary32 formed them in phase B; binary64, at 256 VGPRs, measured
      // faster forming them here)
      if (sizeof(T) == 8 && active && kk + 1 < bt && has1)
        panel_dot2<T, WS, RR>(Lu1, Lu2, Wl, wr, tid, kk, base, pd1, pd2);
      const int npub = 2 + (has1 ? 2 : 0) + (has2 ? 2 : 0);
      const int total = G + npub;
      double sg = 0.0;
      for (int id = tid; id < total; id += NT) {
        int aux;
        if (id < G) {
          const double o = xwait(ar + id, ex, &aux, wg);
          sg = (sg < o) ? o : sg;
        } else {
          const int pid = kPubCr + (id - G);  // 8, 9, (10, 11), (12, 13)
          pubv[pid] = xwait(pubs + pid, ex, &aux, wg);
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
      BBK_PT(8);
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
      if (has1) {
        const T lkk = static_cast<T>(div_or_zero(pubv[kPubWk1], den));
        if (tid < kk) Lnext[tid] = u1;
        if (tid == kk) Lnext[kk] = lkk;
        if (kind == 3) {  // pd1 was formed during exchange 2
          pnext = pd1 + lkk * w;
          have_next = true;
        }
      }
      size = 1;
    } else if (kind == -1) {
      // symmetric interchange k <-> r (src/block_ldlt_decomposition.cc:414-498)
      // columns k and r of the finished panel rows: row q by thread q / G of
      // group q % G, the same thread at every pivot (so its writes of an
      // element follow each other in program order)
      if (qrow < kk) {
        sta(&A[(int64_t)(s + qrow) * lda + k], Lr[qrow]);
        sta(&A[(int64_t)(s + qrow) * lda + r], Lk[qrow]);
      }
      // trailing matrix: new row/column r take the old row k (= arow)
      if (own && j > r)
        sta(&A[(int64_t)r * lda + j], arow);
      else if (own && j > k && j < r)
        sta(&A[(int64_t)j * lda + r], arow);
      if (j == r) sta(&A[(int64_t)r * lda + r], static_cast<T>(pubv[kPubAkk]));
      swap_pos(k, r);
      // the pivot row is now the candidate row: each thread's own value
      const T nw = c;
      if (active) Wl[(kk - base) * WS + tid] = nw;
      const double den = safe_den(crr);
      if (own && j > k)
        sta(&A[(int64_t)k * lda + j], static_cast<T>(div_or_zero(static_cast<double>(nw), den)));
      if (j == k) {
        sta(&A[(int64_t)k * lda + k], nw);
        p.piv[k] = 1;
      }
      if (gi == 0 && tid == 0) {
        swl[2 * nswap] = k;
        swl[2 * nswap + 1] = r;
        p.swaps[2 * nswap] = k;
        p.swaps[2 * nswap + 1] = r;
      }
      ++nswap;
      if (has1) {
        if (r == k + 1) {
          // the new row k+1 is the old row k (the holder of k+1 held k)
          const T lkk = static_cast<T>(div_or_zero(ck, den));
          if (tid < kk) Lnext[tid] = Lk[tid];
          if (tid == kk) Lnext[kk] = lkk;
          arow_next = arow;
          pnext = dotA + lkk * nw;
        } else {
          const T lkk = static_cast<T>(div_or_zero(ck1, den));
          if (tid < kk) Lnext[tid] = u1;
          if (tid == kk) Lnext[kk] = lkk;
          if (j == r) arow_next = static_cast<T>(pubv[kPubAkk1]);
          pnext = pd1 + lkk * nw;
        }
        have_next = true;
      }
      size = 1;
    } else {
      // 2x2 pivot at (k, k+1) after the interchange k+1 <-> r
      // (src/block_ldlt_decomposition.cc:500-587, 658-704)
      const int k1 = k + 1;
      if (r != k1) {
        // columns k+1 and r of the finished panel rows (rows as above)
        if (qrow < kk) {
          sta(&A[(int64_t)(s + qrow) * lda + k1], Lr[qrow]);
          sta(&A[(int64_t)(s + qrow) * lda + r], Lu1[qrow]);
        }
        // trailing matrix: new row/column r take the old row k+1 (= P1)
        if (own && j > r)
          sta(&A[(int64_t)r * lda + j], P1);
        else if (own && j > k1 && j < r)
          sta(&A[(int64_t)j * lda + r], P1);
        if (j == r) sta(&A[(int64_t)r * lda + r], static_cast<T>(a11));
        swap_pos(k1, r);
        if (gi == 0 && tid == 0) {
          swl[2 * nswap] = k1;
          swl[2 * nswap + 1] = r;
          p.swaps[2 * nswap] = k1;
          p.swaps[2 * nswap + 1] = r;
        }
        ++nswap;
      }
      // new W rows kk (current row) and kk+1 (candidate row): each thread's
      // own values (positions k+1 and r exchanged their holders)
      const T w0 = w, w1 = c;
      if (active) Wl[(kk + 1 - base) * WS + tid] = w1;
      const double wkr = pubv[kPubWkr];
      const double a = wkk, bb = wkr, cc = crr;
      double itl, ibr, ioff;
      inverse_2x2_nc(a, bb, cc, &itl, &ibr, &ioff);
      if (own && j >= k + 2) {
        double y0, y1;
        apply_2x2_nc(itl, ibr, ioff, static_cast<double>(w0), static_cast<double>(w1), &y0,
                     &y1);
        sta(&A[(int64_t)k * lda + j], static_cast<T>(y0));
        sta(&A[(int64_t)k1 * lda + j], static_cast<T>(y1));
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
      if (use2) {
        const bool rk2 = (r == k + 2);
        if (tid < kk) Lnext[tid] = rk2 ? u1 : u2;
        const double wa = rk2 ? pubv[kPubWk1] : pubv[kPubWk2];
        const double wb = rk2 ? ck1 : ck2;
        double x0, x1;
        apply_2x2_nc(itl, ibr, ioff, wa, wb, &x0, &x1);
        const T l0 = static_cast<T>(x0), l1 = static_cast<T>(x1);
        if (tid == kk) Lnext[kk] = l0;
        if (tid == kk + 1) Lnext[kk + 1] = l1;
        pnext = (rk2 ? pd1 : pd2) + l0 * w0 + l1 * w1;
        have_next = true;
        if (rk2)
          arow_next = P1;  // the new row k+2 is the old row k+1
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
    BBK_PT(9);
    if (tid == 0 && kind_last != 1) {
      for (int u = 0; u < 9; ++u) pacc[u] += tp[u + 1] - tp[u];
      pacc[9] += 1;
    }
#endif
  }
#if defined(BBK_PANEL_PROFILE)
  if (tid == 0 && (gi == 0 || gi == G - 1))
    for (int i = 0; i < 12; ++i) atomicAdd(&g_panel_profile[(gi == 0 ? 0 : 12) + i], pacc[i]);
#endif

  if (own) {
    if constexpr (RR > 0) {
      if (base > 0) {
#pragma unroll
        for (int q = 0; q < RR; ++q) Wg[(int64_t)q * p.ldw + j] = wr[q];
      }
    }
    for (int q = base; q < kk; ++q) Wg[(int64_t)q * p.ldw + j] = Wl[(q - base) * WS + tid];
  }
  if (gi == 0 && nswap > 0) {
    // the interchanges applied to perm as one gather (kept out of the
    // pivot loop, where a read of perm would stall workgroup 0)
    compose_swaps(swl, nswap, sw_slot, sw_content);
    const int ne = 2 * nswap;
    const int t0 = tid, t1 = tid + NT;
    const int v0 = t0 < ne ? p.perm[sw_content[sw_slot[t0]]] : 0;
    const int v1 = t1 < ne ? p.perm[sw_content[sw_slot[t1]]] : 0;
    wait_vm();
    __syncthreads();  // every read of perm precedes every write
    if (t0 < ne) p.perm[swl[t0]] = v0;
    if (t1 < ne) p.perm[swl[t1]] = v1;
  }
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
hipError_t launch_panel(T* A, const PanelArgs& args, int groups, int nt, hipStream_t stream) {
  T* a = A;
  PanelArgs pa = args;
  void* kargs[] = {&a, &pa};
  const void* fn = nt == kPanelThreadsWide
                       ? reinterpret_cast<const void*>(&bk_panel_kernel<T, kPanelThreadsWide>)
                       : reinterpret_cast<const void*>(&bk_panel_kernel<T, kPanelThreads>);
  return launch_resident(fn, dim3(groups), dim3(nt), kargs, stream, args.status != nullptr);
}

template <typename T>
hipError_t panel_max_groups(int nt, int* groups) {
  int per_cu = 0, dev = 0, cus = 0;
  const void* fn = nt == kPanelThreadsWide
                       ? reinterpret_cast<const void*>(&bk_panel_kernel<T, kPanelThreadsWide>)
                       : reinterpret_cast<const void*>(&bk_panel_kernel<T, kPanelThreads>);
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, nt, 0);
  if (e != hipSuccess) return e;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  *groups = per_cu * cus;
  return hipSuccess;
}

size_t panel_exchange_records(int groups, int elem_bytes) {
  (void)elem_bytes;
  return 2 * static_cast<size_t>(groups + kNumPub);
}

template hipError_t launch_panel<double>(double*, const PanelArgs&, int, int, hipStream_t);
template hipError_t launch_panel<float>(float*, const PanelArgs&, int, int, hipStream_t);
template hipError_t panel_max_groups<double>(int, int*);
template hipError_t panel_max_groups<float>(int, int*);

}  // namespace bbk

#if defined(BBK_PANEL_PROFILE)
// Read (and optionally clear) the phase clock totals of the profiling build.
extern "C" int bbk_panel_profile(unsigned long long* out, int reset) {
  hipError_t e = hipMemcpyFromSymbol(out, HIP_SYMBOL(bbk::g_panel_profile),
                                     sizeof(unsigned long long) * 24);
  if (reset) {
    unsigned long long z[24] = {0};
    (void)hipMemcpyToSymbol(HIP_SYMBOL(bbk::g_panel_profile), z, sizeof(z));
  }
  return e == hipSuccess ? 0 : 1;
}
#endif

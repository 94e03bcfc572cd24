// Shared device-side building blocks for the gfx950 (CDNA4) HODLR kernels:
// MFMA traits and an LDS-staged 64-wide-wavefront tile engine used by the
// batched GEMM, the leaf TRSM and the LU trailing updates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hodlr {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16;

constexpr int kThreads = 256;  // 4 wavefronts of 64 lanes

// --------------------------------------------------------------------- MFMA
// 16x16x4 shapes for fp64 and fp32.  Lane l feeds A(l%16, l/16) and
// B(l/16, l%16) of the 16x4 / 4x16 operand blocks.  Accumulator entry v
// (v = 0..3) of lane l holds D(row(l, v), l%16) with
//   fp64: row = l/16 + 4*v        fp32: row = 4*(l/16) + v
// (both measured on gfx950 with an in-bounds lane probe).
template <class MT> struct Mfma;
template <> struct Mfma<double> {
  typedef f64x4 acc_t;
  typedef double acc_e;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct Mfma<float> {
  typedef f32x4 acc_t;
  typedef float acc_e;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// bf16 16x16x32 (gfx950): lane l feeds A(l%16, 8*(l/16) + e) and
// B(8*(l/16) + e, l%16), e = 0..7; accumulators as fp32 (row 4*(l/16) + v).
template <> struct Mfma<bf16> {
  typedef f32x4 acc_t;
  typedef float acc_e;
  static __device__ __forceinline__ acc_t mma(bf16x8 a, bf16x8 b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// storage -> math type (bf16 rounds to nearest even through fp32)
template <class MT, class TI> __device__ __forceinline__ MT to_math(TI x) { return MT(x); }
template <> __device__ __forceinline__ bf16 to_math<bf16, double>(double x) { return bf16(float(x)); }
template <> __device__ __forceinline__ bf16 to_math<bf16, float>(float x) { return bf16(x); }

// Global-address-space views of device pointers.  Pointers read from
// descriptor tables are generic; through these casts the compiler emits
// global_load/global_store instead of flat ones, which are tracked by vmcnt
// only, so an LDS wait does not also wait for every global load in flight.
#define HODLR_GLOBAL __attribute__((address_space(1)))
template <class T> __device__ __forceinline__ HODLR_GLOBAL T* glob(T* p) { return (HODLR_GLOBAL T*)p; }

// Out-of-range lanes of a staged tile load from here (32 zero bytes): the
// value is already zero, so no select (and no wait for the load) follows the load.
static __device__ __attribute__((aligned(16))) double g_zero_page[4];
template <class T> __device__ __forceinline__ const HODLR_GLOBAL T* zero_page() {
  return (const HODLR_GLOBAL T*)g_zero_page;
}

// Launches of a captured refinement sweep carry a gate: a device flag that the
// convergence check clears once the residual meets the tolerance (or the sweep
// budget is spent), so the sweeps left in the graph exit at their first
// instruction.  A null gate always runs.  The flag is written by an earlier
// kernel of the same stream, whose writes a later kernel sees.
__device__ __forceinline__ bool gated_off(const int32_t* gate) {
  return gate != nullptr && *(const HODLR_GLOBAL int32_t*)gate == 0;
}

// Workgroup barrier that orders LDS only: outstanding global loads stay in
// flight across it (a full __syncthreads() waits for them).  Callers exchange
// data between waves through LDS alone across it; a kernel whose waves read
// global memory other waves of the workgroup wrote uses __syncthreads() after
// those writes.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ------------------------------------------------------------- tile engine
// C (BM x BN) += A (BM x BK) * B (BK x BN) with both operands staged in LDS
// k-major: As[k][m], Bs[k][n].  The four waves split the tile 2 x 2; each wave
// owns (BM/32) x (BN/32) MFMA blocks of 16 x 16.
template <class MT, int BM, int BN, int BK>
struct Tile {
  static constexpr int PA = BM + 1;   // padded LDS row pitch (elements)
  static constexpr int PB = BN + 1;
  static constexpr int WM = BM / 32;
  static constexpr int WN = BN / 32;
  static constexpr int BOFF = BK * PA;           // B stage offset within a stage
  static constexpr int STAGE = BK * (PA + PB);   // one LDS stage (A then B)
  static constexpr int LDS_ELEMS = 2 * STAGE;    // double buffered
  typedef typename Mfma<MT>::acc_t acc_t;
  acc_t acc[WM][WN];

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[i][j][v] = 0;
  }

  // acc += As(BM x KS) * Bs(KS x BN) with pitches pa / pb (KS multiple of 4)
  template <int KS>
  __device__ __forceinline__ void mma(const MT* As, const MT* Bs, int pa = PA, int pb = PB) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = (wave & 1) * (BM / 2), n0 = (wave >> 1) * (BN / 2);
    const int li = lane & 15, lk = lane >> 4;
    // fragments of up to 16 k values are read before their MFMAs so that one
    // LDS wait covers a group of 4 x WM x WN MFMAs
    constexpr int KG = KS < 16 ? KS : 16;
    static_assert(KS % KG == 0 && KG % 4 == 0, "k extent");
#pragma unroll
    for (int k0 = 0; k0 < KS; k0 += KG) {
      MT a[KG / 4][WM], b[KG / 4][WN];
#pragma unroll
      for (int q = 0; q < KG / 4; ++q) {
#pragma unroll
        for (int i = 0; i < WM; ++i) a[q][i] = As[(k0 + 4 * q + lk) * pa + m0 + i * 16 + li];
#pragma unroll
        for (int j = 0; j < WN; ++j) b[q][j] = Bs[(k0 + 4 * q + lk) * pb + n0 + j * 16 + li];
      }
#pragma unroll
      for (int q = 0; q < KG / 4; ++q)
#pragma unroll
        for (int i = 0; i < WM; ++i)
#pragma unroll
          for (int j = 0; j < WN; ++j) acc[i][j] = Mfma<MT>::mma(a[q][i], b[q][j], acc[i][j]);
    }
  }

  // Visit every accumulator entry with its (row, col) inside the tile.
  template <class F> __device__ __forceinline__ void for_each(F&& f) const {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = (wave & 1) * (BM / 2), n0 = (wave >> 1) * (BN / 2);
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          f(m0 + i * 16 + Mfma<MT>::row(lane, v), n0 + j * 16 + (lane & 15), acc[i][j][v]);
  }
};

// bf16 tiles: operands staged "octet-planar", element (k, m) of A at
// As[((k / 8) * BM + m) * 8 + k % 8] (B likewise with n), so each lane reads
// its 8 consecutive k of one row with a single 16-byte LDS load per MFMA.
template <int BM, int BN, int BK>
struct Tile<bf16, BM, BN, BK> {
  static_assert(BK % 32 == 0, "bf16 MFMA consumes k in steps of 32");
  static constexpr int WM = BM / 32;
  static constexpr int WN = BN / 32;
  static constexpr int BOFF = BK * BM;
  static constexpr int STAGE = BK * (BM + BN);
  static constexpr int LDS_ELEMS = 2 * STAGE;
  typedef f32x4 acc_t;
  acc_t acc[WM][WN];

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[i][j][v] = 0;
  }
  template <int KS>
  __device__ __forceinline__ void mma(const bf16* As, const bf16* Bs) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = (wave & 1) * (BM / 2), n0 = (wave >> 1) * (BN / 2);
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int kk = 0; kk < KS; kk += 32) {
      bf16x8 a[WM], b[WN];
#pragma unroll
      for (int i = 0; i < WM; ++i)
        a[i] = *(const bf16x8*)(As + ((kk / 8 + lk) * BM + m0 + i * 16 + li) * 8);
#pragma unroll
      for (int j = 0; j < WN; ++j)
        b[j] = *(const bf16x8*)(Bs + ((kk / 8 + lk) * BN + n0 + j * 16 + li) * 8);
#pragma unroll
      for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) acc[i][j] = Mfma<bf16>::mma(a[i], b[j], acc[i][j]);
    }
  }
  template <class F> __device__ __forceinline__ void for_each(F&& f) const {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = (wave & 1) * (BM / 2), n0 = (wave >> 1) * (BN / 2);
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          f(m0 + i * 16 + Mfma<bf16>::row(lane, v), n0 + j * 16 + (lane & 15), acc[i][j][v]);
  }
};

// Register staging of one A tile (rows [r0, r0+BM) x k [k0, k0+BK)) and one B
// tile (k [k0, k0+BK) x cols [c0, c0+BN)); out-of-range entries are zero.
//   A: TRANS = false -> A(m, k) at A[m + k*lda];  TRANS = true -> A[k + m*lda]
//   B: B(k, n) at B[k + n*ldb]
// Loads are unconditional (an out-of-range lane reads the zero page) and keep
// the storage type: conversion to the math type happens at the LDS store,
// after the MFMA work of the current stage, so no wait sits between a load and
// the MFMAs that run while it is in flight.
// NTB: B loads carry the nontemporal hint (streamed operands that should not
// displace the reused A operand from L2)
template <class MT, class TA, class TB, int BM, int BN, int BK, bool TRANS, bool NTB = false>
struct Stager {
  static constexpr int NA = (BM * BK) / kThreads, NB = (BN * BK) / kThreads;
  static constexpr int PA = BM + 1, PB = BN + 1;
  TA ra[NA];
  TB rb[NB];

  __device__ __forceinline__ void load_a(const HODLR_GLOBAL TA* A, int64_t lda, int r0, int mlim, int k0, int klim) {
#pragma unroll
    for (int s = 0; s < NA; ++s) {
      const int e = threadIdx.x + s * kThreads;
      const int m = TRANS ? e / BK : e % BM, k = TRANS ? e % BK : e / BM;
      const int gm = r0 + m, gk = k0 + k;
      const bool ok = gm < mlim && gk < klim;
      const int64_t off = TRANS ? gk + int64_t(gm) * lda : gm + int64_t(gk) * lda;
      ra[s] = *(ok ? A + off : zero_page<TA>());
    }
  }
  // Interior tiles (every element in range): a wave-uniform base pointer plus
  // a per-thread 32-bit offset that does not change with the k step, so each
  // load needs no address arithmetic and no range check.  Element e = tid +
  // s*kThreads of the tile has inner index e % INNER = tid % INNER and outer
  // index tid / INNER + s * (kThreads / INNER).
  __device__ __forceinline__ void load_a_full(const HODLR_GLOBAL TA* A, int64_t lda, int r0, int k0) {
    constexpr int INNER = TRANS ? BK : BM, STEP = kThreads / INNER;
    static_assert(kThreads % INNER == 0, "tile shape");
    const int tid = threadIdx.x;
    const uint32_t voff = uint32_t(tid % INNER) + uint32_t(tid / INNER) * uint32_t(lda);
    const HODLR_GLOBAL TA* base = TRANS ? A + k0 + int64_t(r0) * lda : A + r0 + int64_t(k0) * lda;
#pragma unroll
    for (int s = 0; s < NA; ++s) ra[s] = (base + int64_t(s) * STEP * lda)[voff];
  }
  __device__ __forceinline__ void load_b_full(const HODLR_GLOBAL TB* B, int64_t ldb, int k0, int c0) {
    constexpr int INNER = BK, STEP = kThreads / INNER;
    const int tid = threadIdx.x;
    const uint32_t voff = uint32_t(tid % INNER) + uint32_t(tid / INNER) * uint32_t(ldb);
    const HODLR_GLOBAL TB* base = B + k0 + int64_t(c0) * ldb;
#pragma unroll
    for (int s = 0; s < NB; ++s) {
      const HODLR_GLOBAL TB* p = &(base + int64_t(s) * STEP * ldb)[voff];
      if constexpr (NTB) rb[s] = __builtin_nontemporal_load(p);
      else rb[s] = *p;
    }
  }
  __device__ __forceinline__ void load(const HODLR_GLOBAL TA* A, int64_t lda, int r0, int mlim,
                                       const HODLR_GLOBAL TB* B, int64_t ldb, int c0, int nlim, int k0, int klim) {
    if (r0 + BM <= mlim && k0 + BK <= klim) load_a_full(A, lda, r0, k0);
    else load_a(A, lda, r0, mlim, k0, klim);
    if (c0 + BN <= nlim && k0 + BK <= klim) load_b_full(B, ldb, k0, c0);
    else load_b(B, ldb, k0, klim, c0, nlim);
  }
  __device__ __forceinline__ void load_b(const HODLR_GLOBAL TB* B, int64_t ldb, int k0, int klim, int c0, int nlim) {
#pragma unroll
    for (int s = 0; s < NB; ++s) {
      const int e = threadIdx.x + s * kThreads;
      const int k = e % BK, n = e / BK;
      const int gk = k0 + k, gn = c0 + n;
      const bool ok = gk < klim && gn < nlim;
      const HODLR_GLOBAL TB* p = ok ? B + gk + int64_t(gn) * ldb : zero_page<TB>();
      if constexpr (NTB) rb[s] = __builtin_nontemporal_load(p);
      else rb[s] = *p;
    }
  }
  __device__ __forceinline__ void store(MT* As, MT* Bs) const {
#pragma unroll
    for (int s = 0; s < NA; ++s) {
      const int e = threadIdx.x + s * kThreads;
      const int m = TRANS ? e / BK : e % BM, k = TRANS ? e % BK : e / BM;
      As[k * PA + m] = to_math<MT>(ra[s]);
    }
#pragma unroll
    for (int s = 0; s < NB; ++s) {
      const int e = threadIdx.x + s * kThreads;
      Bs[(e % BK) * PB + e / BK] = to_math<MT>(rb[s]);
    }
  }
};

// bf16 staging: the same loads, stored octet-planar (see Tile<bf16>)
template <class TA, class TB, int BM, int BN, int BK, bool TRANS, bool NTB>
struct Stager<bf16, TA, TB, BM, BN, BK, TRANS, NTB> : Stager<float, TA, TB, BM, BN, BK, TRANS, NTB> {
  typedef Stager<float, TA, TB, BM, BN, BK, TRANS, NTB> Base;
  using Base::NA;
  using Base::NB;
  __device__ __forceinline__ void store(bf16* As, bf16* Bs) const {
#pragma unroll
    for (int s = 0; s < NA; ++s) {
      const int e = threadIdx.x + s * kThreads;
      const int m = TRANS ? e / BK : e % BM, k = TRANS ? e % BK : e / BM;
      As[((k >> 3) * BM + m) * 8 + (k & 7)] = to_math<bf16>(this->ra[s]);
    }
#pragma unroll
    for (int s = 0; s < NB; ++s) {
      const int e = threadIdx.x + s * kThreads;
      const int k = e % BK, n = e / BK;
      Bs[((k >> 3) * BN + n) * 8 + (k & 7)] = to_math<bf16>(this->rb[s]);
    }
  }
};

// acc += op(A)[r0.., kb..ke) * B[kb..ke), c0..) with register prefetch and two
// LDS stages: global loads of step t+1 are in flight while step t computes.
template <class MT, int BM, int BN, int BK, bool TRANS, bool NTB = false, class TA, class TB>
__device__ __forceinline__ void gemm_loop(Tile<MT, BM, BN, BK>& t, MT* lds, const HODLR_GLOBAL TA* A, int64_t lda,
                                          int r0, int mlim, const HODLR_GLOBAL TB* B, int64_t ldb, int c0, int nlim,
                                          int kb, int ke) {
  typedef Tile<MT, BM, BN, BK> T;
  if (kb >= ke) return;
  Stager<MT, TA, TB, BM, BN, BK, TRANS, NTB> st;
  st.load(A, lda, r0, mlim, B, ldb, c0, nlim, kb, ke);
  st.store(lds, lds + T::BOFF);
  lds_barrier();
  int stage = 0;
  int k0 = kb;
  // steps whose successor is a full k step of an interior tile: branch-free loads
  if (r0 + BM <= mlim && c0 + BN <= nlim) {
    for (; k0 + 2 * BK <= ke; k0 += BK) {
      st.load_a_full(A, lda, r0, k0 + BK);
      st.load_b_full(B, ldb, k0 + BK, c0);
      // the compiler otherwise sinks the loads down to their LDS store, after
      // the MFMAs, and waits for them there: keep them issued ahead of the
      // product and the store (with its wait) behind it
      __builtin_amdgcn_sched_barrier(0);
      MT* cur = lds + stage * T::STAGE;
      t.template mma<BK>(cur, cur + T::BOFF);
      __builtin_amdgcn_sched_barrier(0);
      MT* nxt = lds + (stage ^ 1) * T::STAGE;
      st.store(nxt, nxt + T::BOFF);
      lds_barrier();
      stage ^= 1;
    }
  }
  for (; k0 < ke; k0 += BK) {
    const bool more = k0 + BK < ke;
    if (more) {
      st.load_a(A, lda, r0, mlim, k0 + BK, ke);
      st.load_b(B, ldb, k0 + BK, ke, c0, nlim);
    }
    __builtin_amdgcn_sched_barrier(0);
    MT* cur = lds + stage * T::STAGE;
    t.template mma<BK>(cur, cur + T::BOFF);
    __builtin_amdgcn_sched_barrier(0);
    if (more) {
      MT* nxt = lds + (stage ^ 1) * T::STAGE;
      st.store(nxt, nxt + T::BOFF);
    }
    lds_barrier();
    stage ^= 1;
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs (bid % 8).  Remap so
// that consecutive logical tiles (which share operands) run on one XCD and
// share its L2: logical = (bid % 8) * ceil(n/8) + bid / 8, skipping the holes
// of the last row.
__device__ __forceinline__ int64_t xcd_tile(int64_t bid, int64_t n) {
  if (n < 64) return bid;
  const int64_t per = (n + 7) / 8;           // logical tiles per XCD
  const int64_t x = bid % 8, i = bid / 8;
  const int64_t full = n - (per - 1) * 8;    // XCDs that get `per` tiles
  // XCDs [0, full) own per tiles, the rest per-1
  if (x < full) return x * per + i;
  return full * per + (x - full) * (per - 1) + i;
}

// Index of the descriptor that owns flattened tile `bid` (tile_begin ascending).
template <class D>
__device__ __forceinline__ int find_desc(const D* d, int n, int64_t bid) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (d[mid].tile_begin <= bid) lo = mid; else hi = mid - 1;
  }
  return lo;
}

}  // namespace hodlr

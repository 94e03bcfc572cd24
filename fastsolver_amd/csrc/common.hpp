// Shared device-side building blocks for the gfx950 (CDNA4) HODLR kernels:
// MFMA traits and an LDS-staged 64-wide-wavefront tile engine used by the
// batched GEMM, the leaf TRSM and the LU trailing updates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hodlr {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

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

template <class TO, class TI> __device__ __forceinline__ TO cvt(TI x) { return TO(x); }

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
  static constexpr int LDS_ELEMS = BK * (PA + PB);
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

  __device__ __forceinline__ void mma(const MT* As, const MT* Bs) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = (wave & 1) * (BM / 2), n0 = (wave >> 1) * (BN / 2);
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      MT a[WM], b[WN];
#pragma unroll
      for (int i = 0; i < WM; ++i) a[i] = As[(kk + lk) * PA + m0 + i * 16 + li];
#pragma unroll
      for (int j = 0; j < WN; ++j) b[j] = Bs[(kk + lk) * PB + n0 + j * 16 + li];
#pragma unroll
      for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) acc[i][j] = Mfma<MT>::mma(a[i], b[j], acc[i][j]);
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

// Load an A tile (rows [r0, r0+BM) x k-columns [k0, k0+BK)) into As[k][m].
// TRANS = false: A(m, k) at A[m + k*lda];  TRANS = true: A(m, k) at A[k + m*lda].
// Out-of-range entries (m >= mlim or k >= klim) are zero.
template <class MT, int BM, int BK, bool TRANS, class TA>
__device__ __forceinline__ void load_a(MT* As, const TA* A, int64_t lda, int r0, int mlim, int k0, int klim) {
  constexpr int PA = BM + 1;
  if (!TRANS) {
#pragma unroll
    for (int s = 0; s < (BM * BK) / kThreads; ++s) {
      int e = threadIdx.x + s * kThreads;
      int m = e % BM, k = e / BM;
      int gm = r0 + m, gk = k0 + k;
      MT v = 0;
      if (gm < mlim && gk < klim) v = cvt<MT>(A[gm + int64_t(gk) * lda]);
      As[k * PA + m] = v;
    }
  } else {
#pragma unroll
    for (int s = 0; s < (BM * BK) / kThreads; ++s) {
      int e = threadIdx.x + s * kThreads;
      int k = e % BK, m = e / BK;
      int gm = r0 + m, gk = k0 + k;
      MT v = 0;
      if (gm < mlim && gk < klim) v = cvt<MT>(A[gk + int64_t(gm) * lda]);
      As[k * PA + m] = v;
    }
  }
}

// Load a B tile (k-rows [k0, k0+BK) x columns [c0, c0+BN)) into Bs[k][n];
// B(k, n) at B[k + n*ldb].
template <class MT, int BN, int BK, class TB>
__device__ __forceinline__ void load_b(MT* Bs, const TB* B, int64_t ldb, int k0, int klim, int c0, int nlim) {
  constexpr int PB = BN + 1;
#pragma unroll
  for (int s = 0; s < (BN * BK) / kThreads; ++s) {
    int e = threadIdx.x + s * kThreads;
    int k = e % BK, n = e / BK;
    int gk = k0 + k, gn = c0 + n;
    MT v = 0;
    if (gk < klim && gn < nlim) v = cvt<MT>(B[gk + int64_t(gn) * ldb]);
    Bs[k * PB + n] = v;
  }
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

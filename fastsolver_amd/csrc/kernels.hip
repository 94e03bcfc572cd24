// Batched gfx950 kernels of the HODLR factor-and-solve path.
//
//   gemm_kernel      C = alpha*op(A)*B + beta*C over a descriptor batch, MFMA
//                    16x16x4 (fp64 or fp32), optional deterministic split-K.
//                    Serves the level reduce V^T [d|u] (src/solver/gemm.cc:315-384),
//                    the broadcast d -= u*eta (gemm.cc:417-478), the node-solve
//                    products (src/solver/solver_tasks.cc:236-302), the LU
//                    row solves and trailing updates and the HODLR matvec.
//   lu_*_kernel      batched blocked LU with partial pivoting (getrf semantics)
//                    of the leaf blocks (solver_tasks.cc:437) and of the r x r
//                    Schur complements (solver_tasks.cc:281).
//   trsm_kernel      X_out = U^-1 L^-1 P X_in per (block, 64-column panel),
//                    left-looking, row interchanges applied while gathering X_in.
#include <hip/hip_runtime.h>
#include <math.h>

#include <limits>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace hodlr {

// ======================================================================= GEMM
// ROLE only separates the symbol names of the callers in profiles: 0 node and
// correction products, 1 reduce V^T X, 2 broadcast d -= u eta, 3 LU trailing
// update, 4 the FP64 HODLR matvec of the residuals.
template <int ROLE, class TA, class TB, class TC, class MT, bool TRANS, int BM, int BN>
__global__ __launch_bounds__(kThreads, 3) void gemm_kernel(const GemmDesc* __restrict__ descs, int ndesc,
                                                           const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  constexpr int BK = sizeof(MT) == 2 ? kGemmBKbf16
                   : BN == kGemmNarrowBN ? kGemmNarrowBK
                   : (BM * BN >= 128 * 128 ? kGemmBigBK : (sizeof(MT) == 8 ? kGemmBK : kGemmBK32));
  typedef Tile<MT, BM, BN, BK> TileT;
  typedef typename Mfma<MT>::acc_e AE;
  // the epilogue stages CH columns of the accumulator tile through the same LDS
  constexpr int STAGE_BYTES = TileT::LDS_ELEMS * int(sizeof(MT));
  constexpr int CH = (BM * (BN + 1) * int(sizeof(AE)) <= STAGE_BYTES) ? BN : 32;
  constexpr int PC = BM + 1;
  constexpr int LDS_BYTES = STAGE_BYTES > CH * PC * int(sizeof(AE)) ? STAGE_BYTES : CH * PC * int(sizeof(AE));
  __shared__ __attribute__((aligned(16))) char lds_raw[LDS_BYTES];
  MT* lds = (MT*)lds_raw;

  const int64_t bid = xcd_tile(blockIdx.x, gridDim.x);
  const int p = find_desc(descs, ndesc, bid);
  const GemmDesc d = descs[p];
  const int64_t local = bid - d.tile_begin;
  const int per_split = d.tiles_m * d.tiles_n;
  const int split = int(local / per_split);
  const int rem = int(local % per_split);
  const int r0 = (rem % d.tiles_m) * BM;
  const int c0 = (rem / d.tiles_m) * BN;
  const int kb = split * d.kchunk;
  const int ke = min(d.k, kb + d.kchunk);

  // beta != 0: the C tile is read coalesced (column by column) into registers
  // before the K loop so its latency hides behind the MFMA work (64 x 64 tiles)
  constexpr int CPT = (BM * BN) / kThreads;          // C elements per thread
  constexpr bool PREFETCH = CPT <= 16;
  const bool use_beta = d.nsplit == 1 && d.beta != 0.0;
  const int m = d.m, n = d.n;
  const HODLR_GLOBAL TC* cin = glob((const TC*)d.Cin);
  const int64_t ldi = d.ldcin;
  TC cpre[PREFETCH ? CPT : 1];
  if (PREFETCH && use_beta) {
#pragma unroll
    for (int s = 0; s < (PREFETCH ? CPT : 1); ++s) {
      const int e = threadIdx.x + s * kThreads;
      const int gr = r0 + e % BM, gc = c0 + e / BM;
      cpre[s] = *(gr < m && gc < n ? cin + gr + gc * ldi : zero_page<TC>());
    }
  }

  TileT t;
  t.zero();
  gemm_loop<MT, BM, BN, BK, TRANS>(t, lds, glob((const TA*)d.A), d.lda, r0, d.m, glob((const TB*)d.B), d.ldb, c0, d.n,
                                   kb, ke);

  // epilogue through LDS: CH columns per pass, written column by column
  const double alpha = d.alpha, beta = d.beta;
  HODLR_GLOBAL TC* out = glob((TC*)d.Cout + (d.nsplit > 1 ? int64_t(split) * d.split_stride : 0));
  const int64_t ldo = d.ldcout;
  AE* cs = (AE*)lds_raw;
  constexpr int SPQ = (BM * CH) / kThreads;   // elements per thread and pass
#pragma unroll
  for (int q = 0; q < BN / CH; ++q) {
    t.for_each([&](int r, int c, AE v) {
      if (c / CH == q) cs[(c - q * CH) * PC + r] = v;
    });
    lds_barrier();
#pragma unroll
    for (int s = 0; s < SPQ; ++s) {
      const int e = threadIdx.x + s * kThreads;
      const int r = e % BM, c = e / BM;
      const int gr = r0 + r, gc = c0 + q * CH + c;
      if (gr < m && gc < n) {
        double x = alpha * double(cs[c * PC + r]);
        if (use_beta) {
          // cpre holds the tile column-major in kThreads-strided slots: pass q, slot s
          if constexpr (PREFETCH) x += beta * double(cpre[q * SPQ + s]);
          else x += beta * double(cin[gr + gc * ldi]);
        }
        out[gr + gc * ldo] = TC(x);
      }
    }
    lds_barrier();
  }
}

// ============================================================ row-panel GEMM
// bf16 plans' broadcast d -= u eta (gemm.cc:417-478), C (m x n) = alpha A
// (m x k) B (k x n) + beta C with k <= 128, fp32 storage, bf16 MFMA: one
// workgroup per 64-row panel of C sweeps all n columns in 64-column chunks.
// The panel's rows of A are read once and kept in LDS as MFMA A fragments;
// B (small, L2-resident) goes from global memory straight into the B
// fragments, 32 bytes per lane and k step (B's leading dimension is a
// multiple of 8 and its rows past k are zero); C streams through once, read
// and written coalesced, the next chunk's loads in flight while a chunk
// computes.  HBM traffic: C read + write plus A once.
// Wave w computes columns 16w .. 16w+15 of a chunk for all 64 rows.
template <bool BFRAG>
__global__ __launch_bounds__(kThreads) void panel_gemm_kernel(const GemmDesc* __restrict__ descs, int ndesc,
                                                              const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  constexpr int BM = 64, BN = 64, NKS = 4, PC = BM + 1;   // k <= 128: four 32-k MFMA steps
  constexpr int CPT = BM * BN / kThreads;                 // C elements per thread and chunk
  __shared__ __attribute__((aligned(16))) bf16x8 s_a[4 * NKS * 64];   // [m block][k step][lane]
  __shared__ float s_c[BN * PC];                                      // [col][row]

  const int64_t bid = xcd_tile(blockIdx.x, gridDim.x);
  const GemmDesc d = descs[find_desc(descs, ndesc, bid)];
  const int r0 = int(bid - d.tile_begin) * BM;
  const int m = d.m, n = d.n, k = d.k;
  const int nks = (k + 31) / 32;
  const HODLR_GLOBAL float* A = glob((const float*)d.A);
  const HODLR_GLOBAL float* B = glob((const float*)d.B);
  const HODLR_GLOBAL float* Cin = glob((const float*)d.Cin);
  HODLR_GLOBAL float* Cout = glob((float*)d.Cout);
  const int64_t lda = d.lda, ldb = d.ldb, ldi = d.ldcin, ldo = d.ldcout;
  const float alpha = float(d.alpha), beta = float(d.beta);
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lq = lane >> 4;
  const int fr = tid % BM, fc = __builtin_amdgcn_readfirstlane(tid / BM);   // coalesced: row fr, cols fc + 4q
  const bool rok = r0 + fr < m;

  // A panel (64 rows x k) -> LDS fragments: element (i, kk) at lane
  // (i % 16) + 16 ((kk % 32) / 8), slot kk % 8, of fragment (i / 16, kk / 32)
  {
    bf16* sa = (bf16*)s_a;
    for (int c0 = 0; c0 < NKS * 32; c0 += 32) {
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int kk = c0 + fc + 4 * q;
        v[q] = *(rok && kk < k ? A + (r0 + fr) + int64_t(kk) * lda : zero_page<float>());
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int kk = c0 + fc + 4 * q;
        sa[(((fr / 16) * NKS + kk / 32) * 64 + (fr % 16) + 16 * ((kk % 32) / 8)) * 8 + kk % 8] = bf16(v[q]);
      }
    }
  }
  // the chunk's B fragments: column 16w + li, k = 32 ks + 8 lq .. + 7.
  // Columns past n read column n - 1 and rows past m row m - 1 (their results
  // are never stored), so the chunk loads need no branch.
  // BFRAG: B comes as bf16 MFMA fragments (the FragDesc layout of B, ldb
  // holding its column-group count): one 16-byte load per k step, no rounding
  using BV = std::conditional_t<BFRAG, bf16x8[NKS], f32x4[NKS][2]>;
  const HODLR_GLOBAL f32x4* zb = (const HODLR_GLOBAL f32x4*)zero_page<double>();
  auto load_b = [&](BV& bv, int c0) {
    if constexpr (BFRAG) {
      const int ng = int(ldb);
      const HODLR_GLOBAL bf16x8* fb =
          (const HODLR_GLOBAL bf16x8*)glob((const bf16*)d.B) + int64_t(min(c0 / 16 + wave, ng - 1)) * 64 + lane;
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) bv[ks] = fb[int64_t(min(ks, nks - 1)) * ng * 64];
      return;
    }
    const int col = min(c0 + 16 * wave + li, n - 1);
    const HODLR_GLOBAL f32x4* bc = (const HODLR_GLOBAL f32x4*)(B + int64_t(col) * ldb);
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      const int kk = 32 * ks + 8 * lq;
      const bool ok = ks < nks && kk + 8 <= ldb;
      if constexpr (!BFRAG) {
        bv[ks][0] = *(ok ? bc + kk / 4 : zb);
        bv[ks][1] = *(ok ? bc + kk / 4 + 1 : zb + 1);
      }
    }
  };
  const HODLR_GLOBAL float* crow = Cin + min(r0 + fr, m - 1);
  auto load_c = [&](float (&cv)[CPT], int c0) {
#pragma unroll
    for (int q = 0; q < CPT; ++q) cv[q] = crow[int64_t(min(c0 + fc + 4 * q, n - 1)) * ldi];
  };
  // one chunk: product, accumulators through LDS, C written coalesced
  auto chunk = [&](int c0, const BV& bv, const float (&cv)[CPT]) {
    f32x4 acc[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) acc[mb] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      if (ks < nks) {
        bf16x8 b;
        if constexpr (BFRAG) {
          b = bv[ks];
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            b[e] = bf16(bv[ks][0][e]);
            b[4 + e] = bf16(bv[ks][1][e]);
          }
        }
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
          acc[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(s_a[(mb * NKS + ks) * 64 + lane], b, acc[mb], 0, 0, 0);
      }
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int v = 0; v < 4; ++v) s_c[(16 * wave + li) * PC + 16 * mb + 4 * lq + v] = acc[mb][v];
    lds_barrier();
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const int col = c0 + fc + 4 * q;
      if (rok && col < n) Cout[(r0 + fr) + int64_t(col) * ldo] = alpha * s_c[(fc + 4 * q) * PC + fr] + beta * cv[q];
    }
    lds_barrier();
  };
  // two register sets, no copies between them: the next chunk's loads are
  // issued before a chunk's product and stay in flight through it
  BV bvA, bvB;
  float cvA[CPT], cvB[CPT];
  load_b(bvA, 0);
  load_c(cvA, 0);
  lds_barrier();
  for (int c0 = 0; c0 < n; c0 += 2 * BN) {
    load_b(bvB, c0 + BN);
    load_c(cvB, c0 + BN);
    __builtin_amdgcn_sched_barrier(0);
    chunk(c0, bvA, cvA);
    if (c0 + BN >= n) break;
    load_b(bvA, c0 + 2 * BN);
    load_c(cvA, c0 + 2 * BN);
    __builtin_amdgcn_sched_barrier(0);
    chunk(c0 + BN, bvB, cvB);
  }
}

// ============================================================ FP64 streaming
// C = alpha (A1 B1 + A2 B2) + beta Cin for products with few columns (16 or
// 32 per tile): the residual r = b - (blockdiag(K) x + U Z) as one
// two-segment product per leaf row block (src/solver/gemm.cc:417-478 on c
// columns, in FP64; the matvec's V^T x runs on the tile engine, whose LDS
// tiles read V's columns faster than per-lane transposed loads do).  At
// c <= 32 these carry at most 8 flop per byte of A, under the FP64 ridge
// (~10), so every wave owns 64 rows of C and the tile's columns and streams
// its rows of A once, straight from global memory into FP64 MFMA operand
// registers: no LDS and no barrier in the k loop, the next 16-k step's loads
// in flight while a step computes.  In the e-th MFMA of a step, the k slot q
// of lane (p, q) = (l % 16, l / 16) carries k0 + 4q + e, so a lane's B values
// are 4 consecutive rows of one column:
//   A:  a(b, e) = A(r0 + 16b + p, k0 + 4q + e)   (16 rows of a column per 16 lanes)
//       VEC: a(b, e) = A(r0 + 4p + b, k0 + 4q + e)   (a lane's 4 blocks are 4
//       consecutive rows: one 32-byte load, 512 contiguous bytes per 16 lanes)
//   B:  b(c, e) = B(k0 + 4q + e, 16c + p)
// VEC loads a lane's 4 consecutive values as one 32-byte vector (the host
// selects it when the operands are 32-byte aligned, their leading dimensions
// and m multiples of 4).  Accumulator entry v of lane (p, q), block (b, c):
// C row 16b + q + 4v (VEC: 4(q + 4v) + b), column 16c + p.  The tile goes out
// through LDS, coalesced.
template <int NCB, bool VEC>
__device__ __forceinline__ void stream_segment(f64x4 (&acc)[4][NCB], const HODLR_GLOBAL double* A, int64_t lda,
                                               const HODLR_GLOBAL double* B, int64_t ldb, int k, int r0, int m,
                                               int c0, int n) {
  constexpr int NB = 4, KS = 16;
  constexpr bool QUAD = VEC;
  if (k <= 0) return;
  const int lane = threadIdx.x & 63, p = lane & 15, q = lane >> 4;
  // Rows past m and columns past n read the last row (quad) / column instead
  // (their results are never written), so the loads of a full k step need no
  // predicate.  Each pointer already carries the lane's 4q shift along k.
  const HODLR_GLOBAL double* arow[NB];
  if (QUAD) {
    const int row = min(r0 + 4 * p, m - 4);
    arow[0] = A + row + int64_t(4 * q) * lda;
  } else {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int row = min(r0 + 16 * b + p, m - 1);
      arow[b] = A + row + int64_t(4 * q) * lda;
    }
  }
  const HODLR_GLOBAL double* bcol[NCB];
#pragma unroll
  for (int c = 0; c < NCB; ++c) bcol[c] = B + int64_t(min(c0 + 16 * c + p, n - 1)) * ldb + 4 * q;
  const HODLR_GLOBAL double* zero = zero_page<double>();
  // a full k step: no predicate, no clamp
  auto load = [&](double (&a)[NB][4], double (&bb)[NCB][4], int k0) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t off = int64_t(k0 + e) * lda;   // wave-uniform
      if (QUAD) {
        const f64x4 t = *(const HODLR_GLOBAL f64x4*)(arow[0] + off);
#pragma unroll
        for (int b = 0; b < NB; ++b) a[b][e] = t[b];
      } else {
#pragma unroll
        for (int b = 0; b < NB; ++b) a[b][e] = arow[b][off];
      }
    }
#pragma unroll
    for (int c = 0; c < NCB; ++c) {
      if (VEC) {
        const f64x4 t = *(const HODLR_GLOBAL f64x4*)(bcol[c] + k0);
#pragma unroll
        for (int e = 0; e < 4; ++e) bb[c][e] = t[e];
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) bb[c][e] = bcol[c][k0 + e];
      }
    }
  };
  // the last, partial k step: A at k >= k reads column (row) k - 1, B is 0 there
  auto load_tail = [&](double (&a)[NB][4], double (&bb)[NCB][4], int k0) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int kk = k0 + 4 * q + e;
      const bool kok = kk < k;
      const int dk = (kok ? kk : k - 1) - 4 * q;
      const int64_t off = int64_t(dk) * lda;
#pragma unroll
      for (int b = 0; b < NB; ++b) a[b][e] = QUAD ? arow[0][off + b] : arow[b][off];
#pragma unroll
      for (int c = 0; c < NCB; ++c) bb[c][e] = *(kok ? bcol[c] + (k0 + e) : zero);
    }
  };
  auto compute = [&](const double (&a)[NB][4], const double (&bb)[NCB][4]) {
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int c = 0; c < NCB; ++c) acc[b][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[b][e], bb[c][e], acc[b][c], 0, 0, 0);
  };
  // Full steps two at a time with the next step's loads issued before the
  // current step's products: the loop body is straight-line, so the wait
  // before each product lets the other set's loads stay in flight.
  double a0[NB][4], b0[NCB][4], a1[NB][4], b1[NCB][4];
  const int nfull = k / KS;
  if (nfull > 0) {
    load(a0, b0, 0);
    int s = 0;
    for (; s + 2 < nfull; s += 2) {
      load(a1, b1, (s + 1) * KS);
      __builtin_amdgcn_sched_barrier(0);
      compute(a0, b0);
      __builtin_amdgcn_sched_barrier(0);
      load(a0, b0, (s + 2) * KS);
      __builtin_amdgcn_sched_barrier(0);
      compute(a1, b1);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (s + 1 < nfull) {
      load(a1, b1, (s + 1) * KS);
      __builtin_amdgcn_sched_barrier(0);
      compute(a0, b0);
      compute(a1, b1);
    } else {
      compute(a0, b0);
    }
  }
  if (nfull * KS < k) {
    load_tail(a0, b0, nfull * KS);
    compute(a0, b0);
  }
}

template <int NCB, bool VEC>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void stream_kernel(
    const StreamDesc* __restrict__ descs, int ndesc, const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  constexpr int NB = 4, WR = 64, PC = WR + 1, NC = 16 * NCB;
  __shared__ double s_c[4 * 16 * PC];   // per wave: one 16-column block of its tile
  const int64_t bid = xcd_tile(blockIdx.x, gridDim.x);
  const StreamDesc d = descs[find_desc(descs, ndesc, bid)];
  const int64_t local = bid - d.tile_begin;
  const int tm = int(local % d.tiles_m), tn = int(local / d.tiles_m);
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int p = lane & 15, q = lane >> 4;
  const int r0 = tm * (4 * WR) + wave * WR, c0 = tn * NC;
  const int m = d.m, n = d.n;
  f64x4 acc[NB][NCB];
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int c = 0; c < NCB; ++c) acc[b][c] = f64x4{0, 0, 0, 0};
  if (r0 < m) {
    stream_segment<NCB, VEC>(acc, glob(d.A1), d.lda1, glob(d.B1), d.ldb1, d.k1, r0, m, c0, n);
    if (d.k2 > 0) stream_segment<NCB, VEC>(acc, glob(d.A2), d.lda2, glob(d.B2), d.ldb2, d.k2, r0, m, c0, n);
  }
  // epilogue: each 16-column block of the wave's 64-row tile through the
  // wave's LDS slice, then one row per lane, column by column (coalesced)
  double* sc = s_c + wave * 16 * PC;
  const int row = r0 + lane;
  const double alpha = d.alpha, beta = d.beta;
  const HODLR_GLOBAL double* cin = glob(d.Cin);
  HODLR_GLOBAL double* cout = glob(d.Cout);
#pragma unroll
  for (int c = 0; c < NCB; ++c) {
    const int cb = c0 + 16 * c;
    double old[16];
#pragma unroll
    for (int j = 0; j < 16; ++j)
      old[j] = (beta != 0.0 && row < m && cb + j < n) ? cin[row + int64_t(cb + j) * d.ldcin] : 0.0;
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) sc[p * PC + (VEC ? 4 * (q + 4 * v) + b : 16 * b + q + 4 * v)] = acc[b][c][v];
    __syncthreads();
    if (row < m) {
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (cb + j < n) cout[row + int64_t(cb + j) * d.ldcout] = alpha * sc[j * PC + lane] + beta * old[j];
    }
    __syncthreads();
  }
}

// ========================================================= bf16 level reduce
// The bf16 plans' level reduce W = V_i^T X_i[:, 0:n] (src/solver/gemm.cc:
// 315-384), also on c columns for the correction solves: A = the node's V
// block (k x m) as bf16 MFMA A fragments of V^T (FragDesc layout: a wave's
// fragment is one contiguous 1 KB, and a k step's fragments are adjacent, so
// the loads stay on few pages however long V's columns are), B = the node's
// rows of X in fp32, W in fp64 (or the split-K chunk's partial).  It is HBM-bound -- X in fp32 is most of the
// bytes -- so there is no LDS in the k loop: a workgroup owns 128 rows x 32
// columns of W and one split-K chunk (a multiple of 32 rows), its four waves
// take interleaved 32-row k steps, and each wave loads its MFMA operands
// straight from global memory (16 bytes of V -- 8 consecutive rows of one
// column -- and 32 bytes of X, rounded to bf16, per lane and fragment), the
// next step's loads in flight while a step computes.  The four partial tiles
// are added in LDS in wave order (deterministic).
//   A fragment, block b, lane (p, q): V(k0 + 8q + e, r0 + 16b + p), e < 8
//     (GemmDesc::lda holds the fragment layout's column-group count)
//   B fragment, block c:              X(k0 + 8q + e, c0 + 16c + p)
//   accumulator v of lane (p, q), block (b, c): W(r0 + 16b + 4q + v, c0 + 16c + p)
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void reduce_bf16_kernel(
    const GemmDesc* __restrict__ descs, int ndesc, const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  constexpr int MB = 8, NB = 2, KS = 32, RW = 16 * MB, PC = RW + 1;
  __shared__ float s_w[4 * 16 * PC];   // per wave, one 16-column block of its partial tile
  const int64_t bid = xcd_tile(blockIdx.x, gridDim.x);
  const GemmDesc d = descs[find_desc(descs, ndesc, bid)];
  const int64_t local = bid - d.tile_begin;
  const int per_split = d.tiles_m * d.tiles_n;
  const int split = int(local / per_split), rem = int(local % per_split);
  const int r0 = (rem % d.tiles_m) * RW, c0 = (rem / d.tiles_m) * 16 * NB;
  const int kb = split * d.kchunk, ke = min(d.k, kb + d.kchunk);
  const int m = d.m, n = d.n;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int p = lane & 15, q = lane >> 4;
  // column groups past V's last read its last group, columns past n column
  // n - 1 of X: their results are never written
  const int ng = int(d.lda);
  const HODLR_GLOBAL bf16x8* arow[MB];
#pragma unroll
  for (int b = 0; b < MB; ++b) arow[b] = (const HODLR_GLOBAL bf16x8*)glob((const bf16*)d.A) + min(r0 / 16 + b, ng - 1) * 64 + lane;
  const int64_t kstride = int64_t(ng) * 64;   // bf16x8 slots per 32-row block
  const HODLR_GLOBAL float* bcol[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) bcol[c] = glob((const float*)d.B) + int64_t(min(c0 + 16 * c + p, n - 1)) * d.ldb + 8 * q;
  f32x4 acc[MB][NB];
#pragma unroll
  for (int b = 0; b < MB; ++b)
#pragma unroll
    for (int c = 0; c < NB; ++c) acc[b][c] = f32x4{0, 0, 0, 0};
  auto load = [&](bf16x8 (&a)[MB], float (&bb)[NB][8], int k0) {
#pragma unroll
    for (int b = 0; b < MB; ++b) a[b] = arow[b][(k0 >> 5) * kstride];
#pragma unroll
    for (int c = 0; c < NB; ++c)
#pragma unroll
      for (int e = 0; e < 8; ++e) bb[c][e] = bcol[c][k0 + e];
  };
  // the chunk's last, partial step: rows of X past ke are 0 (the V rows
  // there are finite: the next chunk's rows, or the copy's zero padding)
  auto load_tail = [&](bf16x8 (&a)[MB], float (&bb)[NB][8], int k0) {
#pragma unroll
    for (int b = 0; b < MB; ++b) a[b] = arow[b][(k0 >> 5) * kstride];
#pragma unroll
    for (int c = 0; c < NB; ++c)
#pragma unroll
      for (int e = 0; e < 8; ++e) bb[c][e] = *(k0 + 8 * q + e < ke ? bcol[c] + (k0 + e) : zero_page<float>());
  };
  auto compute = [&](const bf16x8 (&a)[MB], const float (&bb)[NB][8]) {
    bf16x8 bv[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
#pragma unroll
      for (int e = 0; e < 8; ++e) bv[c][e] = bf16(bb[c][e]);
#pragma unroll
    for (int b = 0; b < MB; ++b)
#pragma unroll
      for (int c = 0; c < NB; ++c) acc[b][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[b], bv[c], acc[b][c], 0, 0, 0);
  };
  // this wave's steps: k0 = kb + KS (wave + 4 j), j < cnt; the chunk's
  // partial step, if any, is the last step of one wave
  const int nsteps = ke > kb ? (ke - kb + KS - 1) / KS : 0;
  const int cnt = nsteps > wave ? (nsteps - wave + 3) / 4 : 0;
  const bool tail = cnt > 0 && (ke - kb) % KS != 0 && wave + 4 * (cnt - 1) == nsteps - 1;
  const int nfull = cnt - (tail ? 1 : 0);
  auto k0_of = [&](int j) { return kb + KS * (wave + 4 * j); };
  // full steps with loads issued two steps ahead through three register
  // sets (the straight-line body keeps two sets' loads in flight across each
  // product); loads past the last full step repeat it (never computed)
  bf16x8 a0[MB], a1[MB], a2[MB];
  float b0[NB][8], b1[NB][8], b2[NB][8];
  if (nfull > 0) {
    auto kk = [&](int j) { return k0_of(min(j, nfull - 1)); };
    load(a0, b0, kk(0));
    load(a1, b1, kk(1));
    int j = 0;
    for (; j + 3 <= nfull; j += 3) {
      load(a2, b2, kk(j + 2));
      __builtin_amdgcn_sched_barrier(0);
      compute(a0, b0);
      __builtin_amdgcn_sched_barrier(0);
      load(a0, b0, kk(j + 3));
      __builtin_amdgcn_sched_barrier(0);
      compute(a1, b1);
      __builtin_amdgcn_sched_barrier(0);
      load(a1, b1, kk(j + 4));
      __builtin_amdgcn_sched_barrier(0);
      compute(a2, b2);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (j < nfull) compute(a0, b0);
    if (j + 1 < nfull) compute(a1, b1);
  }
  if (tail) {
    load_tail(a0, b0, k0_of(nfull));
    compute(a0, b0);
  }
  // the four partial tiles added in wave order, 16 columns at a time
  HODLR_GLOBAL double* out = glob((double*)d.Cout + (d.nsplit > 1 ? int64_t(split) * d.split_stride : 0));
  const double alpha = d.alpha;
#pragma unroll
  for (int c = 0; c < NB; ++c) {
#pragma unroll
    for (int b = 0; b < MB; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) s_w[(wave * 16 + p) * PC + 16 * b + 4 * q + v] = acc[b][c][v];
    __syncthreads();
    for (int e = threadIdx.x; e < 16 * RW; e += kThreads) {
      const int row = e % RW, col = e / RW;
      float s = 0.f;
#pragma unroll
      for (int w = 0; w < 4; ++w) s += s_w[(w * 16 + col) * PC + row];
      const int gr = r0 + row, gc = c0 + 16 * c + col;
      if (gr < m && gc < n) out[gr + int64_t(gc) * d.ldcout] = alpha * double(s);
    }
    __syncthreads();
  }
}

// split-K chunks are added in chunk order: deterministic, no atomics
__global__ __launch_bounds__(kThreads) void sum_kernel(const SumDesc* __restrict__ descs, int ndesc,
                                                       const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  const int64_t bid = blockIdx.x;
  const SumDesc d = descs[find_desc(descs, ndesc, bid)];
  const int64_t base = (bid - d.tile_begin) * kSumTile;
  const int64_t total = int64_t(d.m) * d.n;
  const HODLR_GLOBAL double* P = glob((const double*)d.P);
  HODLR_GLOBAL double* out = glob((double*)d.out);
  for (int e = threadIdx.x; e < kSumTile; e += kThreads) {
    int64_t idx = base + e;
    if (idx >= total) break;
    const int64_t i = idx % d.m, j = idx / d.m;
    const HODLR_GLOBAL double* src = P + i + j * d.ldp;
    double s = 0.0;
#pragma unroll 8
    for (int q = 0; q < d.nsplit; ++q) s += src[q * d.stride];
    out[i + j * d.ldo] = s;
  }
}

// ======================================================================= copy
// storage conversion of a copy; bf16 rounds through fp32, as the staged
// bf16 products round their fp64 operands
template <class TD> __device__ __forceinline__ TD store_as(double x) { return TD(x); }
template <> __device__ __forceinline__ bf16 store_as<bf16>(double x) { return bf16(float(x)); }

// V^T as bf16 MFMA A fragments (FragDesc): one thread per 16-byte lane slot
// of kFragPerWave fragments, 8 consecutive rows of one column read (64 bytes)
// and rounded through fp32.  A workgroup writes 4 kFragPerWave fragments of
// one column group and consecutive row blocks, so it reads 16 columns only
// (few pages of src); a wave's loads for all its fragments are issued before
// the first conversion.  VEC: each lane slot is two 32-byte loads (the host
// selects it for src 32-byte aligned, lds a multiple of 4, rows a multiple of
// 8, so a slot is wholly inside or wholly past the rows).  Measured on a
// 2^20 x 1280 V: 2.8 ms (VEC) against 4.0 ms for one fragment per wave of
// 8-byte loads.
template <bool VEC>
__global__ __launch_bounds__(kThreads) void frag_bf16_kernel(const FragDesc* __restrict__ descs, int ndesc,
                                                             const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  constexpr int FPW = kFragPerWave;
  const int64_t bid = blockIdx.x;
  const FragDesc d = descs[find_desc(descs, ndesc, bid)];
  const int64_t t = bid - d.tile_begin;
  const int g = int(t % d.ngroups);
  const int kb0 = int(t / d.ngroups) * (4 * FPW) + (int(threadIdx.x) / 64) * FPW;
  const int nkb = (d.rows + 31) / 32;
  if (kb0 >= nkb) return;
  const int lane = int(threadIdx.x) % 64, p = lane & 15, q = lane >> 4;
  const int col = 16 * g + p;
  const bool cok = col < d.cols;
  const HODLR_GLOBAL double* src = glob(d.src) + int64_t(min(col, d.cols - 1)) * d.lds;
  double v[FPW][8];
#pragma unroll
  for (int f = 0; f < FPW; ++f) {
    const int r0 = 32 * (kb0 + f) + 8 * q;
    if (VEC) {
      const int rc = min(r0, d.rows - 8);
      const f64x4 x0 = *(const HODLR_GLOBAL f64x4*)(src + rc), x1 = *(const HODLR_GLOBAL f64x4*)(src + rc + 4);
      const bool ok = cok && r0 < d.rows;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        v[f][e] = ok ? x0[e] : 0.0;
        v[f][4 + e] = ok ? x1[e] : 0.0;
      }
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[f][e] = *(cok && r0 + e < d.rows ? src + (r0 + e) : zero_page<double>());
    }
  }
  HODLR_GLOBAL bf16x8* dst = (HODLR_GLOBAL bf16x8*)glob((bf16*)d.dst);
#pragma unroll
  for (int f = 0; f < FPW; ++f) {
    const int kb = kb0 + f;
    if (kb >= nkb) break;
    bf16x8 w;
#pragma unroll
    for (int e = 0; e < 8; ++e) w[e] = bf16(float(v[f][e]));
    dst[(int64_t(kb) * d.ngroups + g) * 64 + lane] = w;
  }
}

template <class TS, class TD>
__global__ __launch_bounds__(kThreads) void copy_kernel(const CopyDesc* __restrict__ descs, int ndesc,
                                                        const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  const int64_t bid = blockIdx.x;
  const CopyDesc d = descs[find_desc(descs, ndesc, bid)];
  const int64_t base = (bid - d.tile_begin) * kCopyTile;
  const int64_t total = int64_t(d.rows) * d.cols;
  const HODLR_GLOBAL TS* src = glob((const TS*)d.src);
  HODLR_GLOBAL TD* dst = glob((TD*)d.dst);
  // (row, column) of the tile's first element, divided once per workgroup;
  // an element past the column's end (rows < 2^31 - kCopyTile) wraps with a
  // 32-bit division
  const int64_t j0 = base / d.rows;
  const int i0 = int(base - j0 * d.rows);
  // all of a thread's loads are issued before its stores (past the end: the
  // zero page, not stored)
  constexpr int NE = kCopyTile / kThreads;
  const bool useb = d.beta != 0.0;
  int ii[NE];
  int64_t jj[NE];
  TS v[NE];
  TD w[NE];
#pragma unroll
  for (int u = 0; u < NE; ++u) {
    const int e = threadIdx.x + u * kThreads;
    int i = i0 + e;
    int64_t j = j0;
    if (i >= d.rows) { j += i / d.rows; i %= d.rows; }
    const bool ok = base + e < total;
    ii[u] = i;
    jj[u] = j;
    v[u] = *(ok ? src + i + j * d.lds : zero_page<TS>());
    w[u] = useb ? *(ok ? dst + i + j * d.ldd : zero_page<TD>()) : TD(0);
  }
#pragma unroll
  for (int u = 0; u < NE; ++u) {
    if (base + threadIdx.x + u * kThreads >= total) break;
    double x = d.alpha * double(v[u]);
    if (useb) x += d.beta * double(w[u]);
    if (ii[u] == jj[u]) x += d.diag;
    dst[ii[u] + jj[u] * d.ldd] = store_as<TD>(x);
  }
}

// ========================================================================= LU
// Panel of kLuNB columns starting at k0, factored as kLuNB/PW sub-panels of PW
// columns held in registers: NT = 256 threads, thread t owns rows ks + t +
// q*NT (q < RPT) of the sub-panel, RPT = 1, 2 or 4 (the smallest that covers
// the batch's largest block).  A 256-thread group keeps the column steps'
// barriers cheap and lets two blocks share a CU, so one block's pivot latency
// hides behind the other's.  Panels of at most 256 rows use one wave per
// block (NT = 64).  PW = 16 (fp64, or fp32 with RPT = 4) or 32 (fp32).
// Partial pivoting picks the first largest |a| (idamax); the column below a
// zero pivot is left unscaled and info records it, as dgetf2 does.
constexpr int kPanelNT = 256;
constexpr int kPanelThreads = 4 * kPanelNT;   // largest block the panel kernel holds

// sub-panel width: a[RPT][PW] stays within 128 registers
template <class T, int RPT> struct PanelWidth { static constexpr int value = sizeof(T) == 4 && RPT < 4 ? 32 : 16; };
// waves per SIMD the compiler is held to: fp32 panels of 769-1024-row blocks
// fit 128 registers (a few spills), so four blocks share a CU instead of two
// (the panels of a batch of 1024 leaves of 1024 rows are all resident at once:
// 5.7 -> 4.7 ms per factorization); used for batches of more blocks than two
// per CU (256 blocks, one per CU, ran 5 % slower with the spills); the other
// variants keep the compiler's register choice (fp64 at 168 registers spills
// heavily)
template <class T, int NT, int RPT> struct PanelOcc {
  static constexpr int value = sizeof(T) == 4 && NT == 256 && RPT == 4 ? 4 : 1;
};
constexpr int kPanelOccBlocks = 512;   // two blocks per CU on 256 CUs

template <class T>
__device__ __forceinline__ void pivot_combine(T& bv, int& bi, T ov, int oi) {
  if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
}

// (value, index) argmax over the 64 lanes of a wave with DPP row shifts and
// row broadcasts (VALU latency) instead of LDS-routed shuffles: after
// row_shr 1/2/4/8 lane 15 of each row holds its row's best, row_bcast 15/31
// carry it to lane 63, which every lane then reads.  Ties keep the smaller
// index, as pivot_combine does.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i32(int x) { return __builtin_amdgcn_update_dpp(x, x, CTRL, ROWS, 0xf, false); }
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_val(float x) { return __int_as_float(dpp_i32<CTRL, ROWS>(__float_as_int(x))); }
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_val(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_i32<CTRL, ROWS>(int(b)), hi = dpp_i32<CTRL, ROWS>(int(b >> 32));
  return __longlong_as_double((long long)(unsigned(lo)) | ((long long)hi << 32));
}
template <class T, int CTRL, int ROWS>
__device__ __forceinline__ void argmax_step(T& bv, int& bi) {
  const T ov = dpp_val<CTRL, ROWS>(bv);
  const int oi = dpp_i32<CTRL, ROWS>(bi);
  pivot_combine(bv, bi, ov, oi);
}
template <class T>
__device__ __forceinline__ void wave_argmax(T& bv, int& bi) {
  argmax_step<T, 0x111, 0xf>(bv, bi);   // row_shr:1
  argmax_step<T, 0x112, 0xf>(bv, bi);   // row_shr:2
  argmax_step<T, 0x114, 0xf>(bv, bi);   // row_shr:4
  argmax_step<T, 0x118, 0xf>(bv, bi);   // row_shr:8
  argmax_step<T, 0x142, 0xa>(bv, bi);   // row_bcast:15 into rows 1 and 3
  argmax_step<T, 0x143, 0xc>(bv, bi);   // row_bcast:31 into rows 2 and 3
  if constexpr (sizeof(T) == 8) {
    const long long b = __double_as_longlong(bv);
    const int lo = __builtin_amdgcn_readlane(int(b), 63), hi = __builtin_amdgcn_readlane(int(b >> 32), 63);
    bv = __longlong_as_double((long long)(unsigned(lo)) | ((long long)hi << 32));
  } else {
    bv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(bv), 63));
  }
  bi = __builtin_amdgcn_readlane(bi, 63);
}

template <class PT>
__device__ __forceinline__ void swap_rows_in_column(PT A, int64_t col, int r1, int r2) {
  if (r1 != r2) {
    const auto x = A[r1 + col];
    A[r1 + col] = A[r2 + col];
    A[r2 + col] = x;
  }
}

template <class T, int NT, int RPT, int OCC = 1>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(OCC))) void lu_panel_kernel(
    const LuDesc* __restrict__ descs, int k0) {
  constexpr int NW = NT / 64;
  constexpr int PW = PanelWidth<T, RPT>::value;
  constexpr int NSP = kLuNB / PW;
  constexpr int PX = kLuNB + 1;
  static_assert(PW <= NT, "row j of a sub-panel is row j of thread j");
  const LuDesc d = descs[blockIdx.x];
  const int m = d.m;
  if (k0 >= m) return;
  HODLR_GLOBAL T* A = glob((T*)d.A);
  const int64_t lda = d.lda;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nb = min(kLuNB, m - k0);

  __shared__ T s_rowp[2][PW];
  __shared__ T s_rowj[2][PW];
  __shared__ T s_rv[2][NW];
  __shared__ int s_ri[2][NW];
  __shared__ int s_piv[kLuNB];        // absolute pivot row of panel column jg
  __shared__ T s_l[PW * (PW + 1)];    // L11 of the current sub-panel
  __shared__ T s_x[PW * PX];          // U12 rows of the current sub-panel
  __shared__ int s_info;
  if (tid == 0) s_info = 0;
  lds_barrier();

  for (int sp = 0; sp < NSP; ++sp) {
    const int ks = k0 + PW * sp;
    if (ks >= k0 + nb) break;
    const int w = min(PW, k0 + nb - ks);
    const int rows = m - ks;
    T a[RPT][PW];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = tid + q * NT;
#pragma unroll
      for (int jj = 0; jj < PW; ++jj) a[q][jj] = (r < rows && jj < w) ? A[(ks + r) + int64_t(ks + jj) * lda] : T(0);
    }
#pragma unroll
    for (int j = 0; j < PW; ++j) {
      if (j < w) {
        const int buf = j & 1;
        // first largest |a| of the thread's rows at or below j (NaN never wins)
        T bv = T(-1);
        int bi = 0x7fffffff;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          const int r = tid + q * NT;
          const T v = fabs(a[q][j]);
          if (r < rows && r >= j && v > bv) { bv = v; bi = r; }
        }
        wave_argmax(bv, bi);
        if (lane == 0) { s_rv[buf][wave] = bv; s_ri[buf][wave] = bi; }
        lds_barrier();
        T pv = s_rv[buf][0];
        int p = s_ri[buf][0];
#pragma unroll
        for (int q = 1; q < NW; ++q) pivot_combine(pv, p, s_rv[buf][q], s_ri[buf][q]);
        if (p >= rows) p = j;   // no candidate (all NaN): keep row j
        // (row tests rather than a computed q keep a[][] statically indexed)
#pragma unroll
        for (int q = 0; q < RPT; ++q)
          if (tid + q * NT == p) {
#pragma unroll
            for (int jj = 0; jj < PW; ++jj) s_rowp[buf][jj] = a[q][jj];
          }
        if (tid == j) {
#pragma unroll
          for (int jj = 0; jj < PW; ++jj) s_rowj[buf][jj] = a[0][jj];
        }
        if (tid == 0) {
          s_piv[PW * sp + j] = ks + p;
          d.ipiv[ks + j] = ks + p;
          if (!(pv > T(0)) && s_info == 0) s_info = ks + j + 1;
        }
        lds_barrier();
        if (p != j) {
          if (tid == j) {
#pragma unroll
            for (int jj = 0; jj < PW; ++jj) a[0][jj] = s_rowp[buf][jj];
          }
#pragma unroll
          for (int q = 0; q < RPT; ++q)
            if (tid + q * NT == p) {
#pragma unroll
              for (int jj = 0; jj < PW; ++jj) a[q][jj] = s_rowj[buf][jj];
            }
        }
        const T piv = s_rowp[buf][j];
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          const int r = tid + q * NT;
          if (r < rows && r > j) {
            const T l = piv != T(0) ? a[q][j] / piv : a[q][j];
            a[q][j] = l;
#pragma unroll
            for (int jj = 0; jj < PW; ++jj)   // fixed trip count: unrolled before the j loop
              if (jj > j) a[q][jj] -= l * s_rowp[buf][jj];
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = tid + q * NT;
      if (r < rows) {
#pragma unroll
        for (int jj = 0; jj < PW; ++jj)
          if (jj < w) A[(ks + r) + int64_t(ks + jj) * lda] = a[q][jj];
      }
    }
    if (tid < PW) {
#pragma unroll
      for (int jj = 0; jj < PW; ++jj) s_l[tid * (PW + 1) + jj] = a[0][jj];
    }
    lds_barrier();
    // this sub-panel's interchanges on the other panel columns, in order
    const int nleft = ks - k0, nright = k0 + nb - (ks + PW);
    for (int t = tid; t < nleft + max(nright, 0); t += NT) {
      const int col = t < nleft ? k0 + t : ks + PW + (t - nleft);
      const int64_t cb = int64_t(col) * lda;
      for (int j = 0; j < w; ++j) swap_rows_in_column(A, cb, ks + j, s_piv[PW * sp + j]);
    }
    __syncthreads();
    if (nright > 0) {
      // 4 loads in flight per thread before their LDS writes
      for (int e0 = tid; e0 < PW * nright; e0 += 4 * NT) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = e0 + u * NT;
          v[u] = *(e < PW * nright ? A + (ks + e % PW) + int64_t(ks + PW + e / PW) * lda : zero_page<T>());
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = e0 + u * NT;
          if (e < PW * nright) s_x[(e % PW) * PX + e / PW] = v[u];
        }
      }
      lds_barrier();
      // U12 = L11^-1 A12 (unit lower), one thread per column
      if (tid < nright) {
        for (int i = 1; i < PW; ++i) {
          T x = s_x[i * PX + tid];
          for (int k = 0; k < i; ++k) x -= s_l[i * (PW + 1) + k] * s_x[k * PX + tid];
          s_x[i * PX + tid] = x;
        }
      }
      lds_barrier();
      for (int e = tid; e < PW * nright; e += NT) {
        const int r = e % PW, c = e / PW;
        A[(ks + r) + int64_t(ks + PW + c) * lda] = s_x[r * PX + c];
      }
      // A22 -= L21 U12 on the remaining panel columns, L21 rows from registers.
      // The column loads run PD columns ahead of the updates (a column past
      // nright reloads the last one and is never stored), so a thread waits for
      // memory about once per PD columns rather than once per column.
      constexpr int PD = sizeof(T) == 8 && RPT == 4 ? 2 : 4;
      T pre[PD][RPT];
      auto load_col = [&](T* dst, int c) {
        const int64_t cb = int64_t(ks + PW + min(c, nright - 1)) * lda;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          const int r = tid + q * NT;
          dst[q] = (r >= PW && r < rows) ? A[(ks + r) + cb] : T(0);
        }
      };
#pragma unroll
      for (int u = 0; u < PD; ++u) load_col(pre[u], u);
      for (int c0 = 0; c0 < nright; c0 += PD) {
#pragma unroll
        for (int u = 0; u < PD; ++u) {
          const int c = c0 + u;
          T x[RPT];
#pragma unroll
          for (int q = 0; q < RPT; ++q) x[q] = pre[u][q];
          load_col(pre[u], c + PD);
          if (c < nright) {
#pragma unroll
            for (int k = 0; k < PW; ++k) {
              const T uv = s_x[k * PX + c];
#pragma unroll
              for (int q = 0; q < RPT; ++q) x[q] -= a[q][k] * uv;
            }
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
              const int r = tid + q * NT;
              if (r >= PW && r < rows) A[(ks + r) + int64_t(ks + PW + c) * lda] = x[q];
            }
          }
        }
      }
      __syncthreads();
    }
  }
  // panel interchanges on every column outside the panel (dlaswp); nothing
  // to do when the panel kept its rows
  bool any = false;
  for (int j = 0; j < nb; ++j) any |= s_piv[j] != k0 + j;
  if (any) {
    const int nout = m - nb;
    for (int c = tid; c < nout; c += NT) {
      const int col = c < k0 ? c : c + nb;
      const int64_t cb = int64_t(col) * lda;
      for (int j = 0; j < nb; ++j) swap_rows_in_column(A, cb, k0 + j, s_piv[j]);
    }
  }
  if (tid == 0 && s_info != 0 && *d.info == 0) *d.info = s_info;
}

// f(integral_constant<I>) for I = B .. E-1, each call with a compile-time index
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>());
    static_for<B + 1, E>(f);
  }
}

// Whole LU with partial pivoting of blocks of at most G * NU rows (the node
// systems S, small leaves), one G x G-thread workgroup per block (G = 8, 16
// or 32), the block in registers: thread (ty, tx) = (t / G, t % G) holds
// A(ty + Gu, tx + Gw), u, w < NU (zero past m), so a column step costs each
// thread at most NU^2 FMAs.  Column step j (its column block v = j / G is a
// compile-time index, so blocks above / left of the step are skipped
// statically):
//   1. the owners of column j post their best candidate (first largest |a|
//      at or below row j: idamax, NaN never wins); lanes 0..G-1 of every
//      wave read the G candidates and combine them with DPP row shifts (and,
//      for G = 32, a row broadcast);
//   2. the owners of rows j and p post them; the exchange is done in
//      registers; the owners of column j scale it by the pivot's reciprocal
//      (dgetf2's dscal; unscaled below a zero pivot, whose column info
//      records) and post it;
//   3. every thread applies the rank-1 update to its entries right of and
//      below the pivot (the posted L column is 0 at rows <= j, the U row 0 at
//      columns <= j, so the update needs no predicate).
// Three LDS-only barriers per column, no global memory between the load and
// the store of the block.  Writes LU in place, ipiv and info; the inverted
// diagonal blocks and perm come from lu_dinv_kernel and lu_perm_kernel.
template <class T, int G, int NU>
__global__ __launch_bounds__(G * G) void lu_small_kernel(const LuDesc* __restrict__ descs) {
  static_assert(G == 8 || G == 16 || G == 32, "thread grid side");
  constexpr int MAXM = G * NU;   // largest block
  const LuDesc d = descs[blockIdx.x];
  const int m = d.m;
  if (m <= 0) return;
  HODLR_GLOBAL T* A = glob((T*)d.A);
  const int64_t lda = d.lda;
  const int tid = threadIdx.x, tx = tid % G, ty = tid / G, lane = tid & 63;
  __shared__ T s_rj[MAXM], s_rp[MAXM], s_l[MAXM];
  __shared__ T s_cv[G];
  __shared__ int s_ci[G];
  __shared__ int s_ipiv[MAXM];
  T a[NU][NU];
#pragma unroll
  for (int u = 0; u < NU; ++u)
#pragma unroll
    for (int w = 0; w < NU; ++w) {
      const int r = ty + G * u, c = tx + G * w;
      a[u][w] = *(r < m && c < m ? A + r + int64_t(c) * lda : zero_page<T>());
    }
  int info = 0;
  // column block v as a template constant (a loop the compiler might not
  // unroll would index a[][] at run time, through scratch memory)
  auto block = [&](auto vc) __attribute__((always_inline)) {
    constexpr int v = decltype(vc)::value;
    if (G * v >= m) return;
    for (int tj = 0; tj < G && G * v + tj < m; ++tj) {
      const int j = G * v + tj;
      if (tx == tj) {
        T bv = T(-1);
        int bi = 0x7fffffff;
#pragma unroll
        for (int u = v; u < NU; ++u) {
          const int r = ty + G * u;
          const T x = fabs(a[u][v]);
          if (r >= j && r < m && x > bv) { bv = x; bi = r; }
        }
        s_cv[ty] = bv;
        s_ci[ty] = bi;
      }
      lds_barrier();
      T pv = s_cv[lane & (G - 1)];
      int p = s_ci[lane & (G - 1)];
      argmax_step<T, 0x111, 0xf>(pv, p);   // row_shr:1
      argmax_step<T, 0x112, 0xf>(pv, p);   // row_shr:2
      argmax_step<T, 0x114, 0xf>(pv, p);   // row_shr:4: lane 7 holds the best of lanes 0..7
      // row_shr:8: lane 15 of each 16-lane row holds its best
      if constexpr (G >= 16) argmax_step<T, 0x118, 0xf>(pv, p);
      // row_bcast:15 into rows 1 and 3: lane 31 holds the best of 32
      if constexpr (G == 32) argmax_step<T, 0x142, 0xa>(pv, p);
      p = __builtin_amdgcn_readlane(p, G - 1);
      if constexpr (sizeof(T) == 8) {
        const long long bits = __double_as_longlong(pv);
        const int lo = __builtin_amdgcn_readlane(int(bits), G - 1);
        const int hi = __builtin_amdgcn_readlane(int(bits >> 32), G - 1);
        pv = __longlong_as_double((long long)(unsigned(lo)) | ((long long)hi << 32));
      } else {
        pv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pv), G - 1));
      }
      if (p >= m) p = j;   // no candidate (all NaN): keep row j
      if (!(pv > T(0)) && info == 0) info = j + 1;
      if (tid == 0) s_ipiv[j] = p;
      const int pu = p / G, pt = p % G;
      if (ty == tj && p != j) {
#pragma unroll
        for (int w = 0; w < NU; ++w) s_rj[tx + G * w] = a[v][w];
      }
      if (ty == pt) {
#pragma unroll
        for (int u = v; u < NU; ++u) {
          // (an opaque test per block: the compiler would otherwise fold the
          // chain into a[pu][w], a run-time index into scratch memory)
          int pick = u == pu;
          asm volatile("" : "+v"(pick));
          if (pick) {
#pragma unroll
            for (int w = 0; w < NU; ++w) s_rp[tx + G * w] = a[u][w];
          }
        }
      }
      lds_barrier();
      if (p != j) {
        if (ty == tj) {
#pragma unroll
          for (int w = 0; w < NU; ++w) a[v][w] = s_rp[tx + G * w];
        }
        if (ty == pt) {
#pragma unroll
          for (int u = v; u < NU; ++u) {
            int pick = u == pu;
            asm volatile("" : "+v"(pick));
            if (pick) {
#pragma unroll
              for (int w = 0; w < NU; ++w) a[u][w] = s_rj[tx + G * w];
            }
          }
        }
      }
      if (tx == tj) {
        const T piv = s_rp[j];
        const T rcp = piv != T(0) ? T(1) / piv : T(1);
#pragma unroll
        for (int u = v; u < NU; ++u) {
          const int r = ty + G * u;
          T l = T(0);
          if (r > j && r < m) {
            l = a[u][v] * rcp;
            a[u][v] = l;
          }
          s_l[r] = l;
        }
      }
      lds_barrier();
#pragma unroll
      for (int u = v; u < NU; ++u) {
        const T lu = s_l[ty + G * u];
#pragma unroll
        for (int w = v; w < NU; ++w) a[u][w] -= lu * (tx + G * w > j ? s_rp[tx + G * w] : T(0));
      }
    }
  };
  static_for<0, NU>(block);
#pragma unroll
  for (int u = 0; u < NU; ++u)
#pragma unroll
    for (int w = 0; w < NU; ++w) {
      const int r = ty + G * u, c = tx + G * w;
      if (r < m && c < m) A[r + int64_t(c) * lda] = a[u][w];
    }
  for (int j = tid; j < m; j += G * G) d.ipiv[j] = s_ipiv[j];   // (the last barrier ordered s_ipiv)
  if (tid == 0) *d.info = info;
}

// perm from the sequential interchanges ipiv (row i of P*A is row perm[i] of A)
constexpr int kPermMax = 8192;
__global__ __launch_bounds__(64) void lu_perm_kernel(const LuDesc* __restrict__ descs) {
  const LuDesc d = descs[blockIdx.x];
  const int m = d.m;
  if (m <= kPermMax) {
    __shared__ int s_p[kPermMax];
    for (int i = threadIdx.x; i < m; i += 64) s_p[i] = i;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 0; i < m; ++i) {
        const int q = d.ipiv[i];
        const int x = s_p[i];
        s_p[i] = s_p[q];
        s_p[q] = x;
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += 64) d.perm[i] = s_p[i];
  } else {
    if (threadIdx.x == 0) {
      for (int i = 0; i < m; ++i) d.perm[i] = i;
      for (int i = 0; i < m; ++i) {
        const int q = d.ipiv[i];
        const int x = d.perm[i];
        d.perm[i] = d.perm[q];
        d.perm[q] = x;
      }
    }
  }
}

__global__ void lu_init_kernel(const LuDesc* __restrict__ descs, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) *descs[i].info = 0;
}

// ====================================================== diagonal-block inverses
// For the 64-row diagonal block b0 + blockIdx.x of every LU factor: L_bb^-1
// (unit lower) and U_bb^-1 (upper), zero outside the block's nr x nr corner.
// The block is final once its panel is factored, so the plan inverts it right
// after the panel: the solve for the U rows, U12 = L11^-1 A12, and the leaf TRSM
// then run as MFMA products.  Wave 0 builds column j = lane of L^-1, wave 1
// column j of U^-1, each column held in registers and updated column-oriented
// (x_i -= A_ik x_k for all i at once), so the 64-step chains carry up to 63
// independent FMAs per step and the block entries are wave-uniform LDS reads.
constexpr int kDB = 64;

template <class T>
__global__ __launch_bounds__(128) void lu_dinv_kernel(const LuDesc* __restrict__ descs, int b0) {
  const LuDesc d = descs[blockIdx.y];
  const int b = b0 + blockIdx.x;
  const int r0 = b * kDB;
  if (r0 >= d.m) return;
  const int nr = min(kDB, d.m - r0);
  const HODLR_GLOBAL T* A = glob((const T*)d.A);
  // s_a[k*kDB + i] = A(r0+i, r0+k); the identity outside the nr x nr corner
  __shared__ __attribute__((aligned(16))) T s_a[kDB * kDB];
  {
    // thread t reads row t % kDB of columns t / kDB + 2q, 8 columns' loads in
    // flight before their LDS writes
    const int i = threadIdx.x % kDB, w = __builtin_amdgcn_readfirstlane(threadIdx.x / kDB);
#pragma unroll
    for (int q0 = 0; q0 < kDB / 2; q0 += 8) {
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = w + 2 * (q0 + u);
        const HODLR_GLOBAL T* col = A + int64_t(k + r0) * d.lda;
        v[u] = *(i < nr && k < nr ? col + (r0 + i) : zero_page<T>());
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = w + 2 * (q0 + u);
        s_a[k * kDB + i] = (i < nr && k < nr) ? v[u] : (i == k ? T(1) : T(0));
      }
    }
  }
  __syncthreads();
  const int j = threadIdx.x & 63;
  T x[kDB];
#pragma unroll
  for (int i = 0; i < kDB; ++i) x[i] = i == j ? T(1) : T(0);
  if (threadIdx.x < 64) {
    // L x = e_j, forward: x_k is final at step k
#pragma unroll
    for (int k = 0; k < kDB - 1; ++k) {
#pragma unroll
      for (int i = k + 1; i < kDB; ++i) {
        x[i] -= s_a[k * kDB + i] * x[k];
        if ((i - k) % 16 == 0) __builtin_amdgcn_sched_barrier(0);   // at most 16 entries in flight
      }
    }
  } else {
    // U x = e_j, backward; x_k stays 0 for k > j, where the divisor is 1 so a
    // zero pivot outside the column's range does not turn it into NaN
#pragma unroll
    for (int k = kDB - 1; k >= 0; --k) {
      x[k] = x[k] / (k <= j ? s_a[k * kDB + k] : T(1));
#pragma unroll
      for (int i = 0; i < k; ++i) {
        x[i] -= s_a[k * kDB + i] * x[k];
        if (i % 16 == 15) __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  // column j of the inverse, contiguous in the column-major output (ld kDB)
  HODLR_GLOBAL T* out = glob((T*)d.dinv + int64_t(b) * 2 * kDB * kDB + (threadIdx.x < 64 ? 0 : kDB * kDB) +
                             int64_t(j) * kDB);
  const bool live = j < nr;
#pragma unroll
  for (int i = 0; i < kDB; ++i) out[i] = live ? x[i] : T(0);
}

// ======================================================================= TRSM
// One workgroup per (block, 64-column panel): Y = L^-1 P X_in, then
// X_out = U^-1 Y, 64 rows at a time.  Each step applies the rows already
// solved with a prefetching MFMA product, then multiplies by the inverted
// diagonal block (MFMA on an LDS-resident 64 x 64 pair), so no step runs a
// serial substitution.
template <class TLU, class TIN, class TOUT, class UT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3))) void trsm_kernel(
    const TrsmDesc* __restrict__ descs, int ndesc, const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  // MT: type of the diagonal-block products and of the stored factor; UT: math
  // type of the products with the solved rows (MT, or bf16 in the bf16 plans)
  typedef TLU MT;
  constexpr int BK = sizeof(MT) == 8 ? kTrsmBK : kTrsmBK32, RB = kTrsmRB, CB = kTrsmCB, P = RB + 1;
  constexpr int PT = CB + 1;   // pitch of the right-hand-side block
  constexpr int BKU = sizeof(UT) == 2 ? kGemmBKbf16 : BK;
  static_assert(RB == kDB, "diagonal blocks and TRSM steps share the 64-row blocking");
  typedef Tile<MT, RB, CB, BK> TileT;
  typedef Tile<UT, RB, CB, BKU> TileU;
  constexpr int UBYTES = TileU::LDS_ELEMS * int(sizeof(UT)), DBYTES = RB * P * int(sizeof(MT));
  constexpr int TBYTES = RB * PT * int(sizeof(MT)), HK = RB / 2, HBYTES = HK * P * int(sizeof(MT));
  constexpr int UDB = UBYTES > DBYTES ? UBYTES : DBYTES;
  // fp64: the right-hand-side block reuses the MFMA stages (it is filled after
  // the update loop) and the inverted diagonal block is applied in two k
  // halves, which keeps a workgroup under 53 KB of LDS (3 per CU instead of 2)
  constexpr bool COMPACT = sizeof(MT) == 8;
  constexpr int LBYTES = COMPACT ? (UBYTES > TBYTES + HBYTES ? UBYTES : TBYTES + HBYTES) : UDB + TBYTES;
  __shared__ __attribute__((aligned(16))) char lds_raw[LBYTES];
  UT* s_stage = (UT*)lds_raw;                                  // MFMA stages of the update
  MT* s_t = (MT*)(lds_raw + (COMPACT ? 0 : UDB));            // right-hand-side block, s_t[r*PT + c]
  MT* s_u = (MT*)(lds_raw + (COMPACT ? TBYTES : 0));         // inverted diagonal block (As layout)

  const int64_t bid = xcd_tile(blockIdx.x, gridDim.x);
  const TrsmDesc d = descs[find_desc(descs, ndesc, bid)];
  const int c0 = int(bid - d.tile_begin) * CB;
  const int m = d.m, ncols = d.ncols;
  const int nc = min(CB, ncols - c0);
  const HODLR_GLOBAL TLU* LU = glob((const TLU*)d.LU);
  const HODLR_GLOBAL TLU* DI = glob((const TLU*)d.dinv);
  const HODLR_GLOBAL TIN* Xin = glob((const TIN*)d.Xin);
  HODLR_GLOBAL TOUT* Xout = glob((TOUT*)d.Xout);
  const HODLR_GLOBAL int32_t* perm = d.perm ? glob(d.perm) : nullptr;
  const int64_t ldlu = d.ldlu, ldin = d.ldin, ldout = d.ldout;
  const int nblk = (m + RB - 1) / RB;
  // fp64: the solved rows are read with the nontemporal hint, which keeps
  // more of the factor's block rows (shared by the leaf's column panels) in
  // L2 (measured 66.0 -> 64.4 ms for 1024 leaves of 1024 x 1312; slower for
  // fp32 and bf16)
  constexpr bool NTY = sizeof(UT) == 8;
  TileT t;
  TileU tu;

  for (int pass = 0; pass < 2; ++pass) {        // 0: forward with L, 1: backward with U
    for (int s = 0; s < nblk; ++s) {
      const int b = pass == 0 ? s : nblk - 1 - s;
      const int r0 = b * RB, nr = min(RB, m - r0);
      tu.zero();
      if (pass == 0)
        gemm_loop<UT, RB, CB, BKU, false, NTY>(tu, s_stage, LU, ldlu, r0, m, Xout, ldout, c0, ncols, 0, r0);
      else
        gemm_loop<UT, RB, CB, BKU, false, NTY>(tu, s_stage, LU, ldlu, r0, r0 + nr, Xout, ldout, c0, ncols, r0 + RB, m);
      // Right-hand-side rows of the step into s_t, read coalesced (thread t
      // owns row t % RB of columns t / RB + q * CS, 64 rows of a column per
      // wave), and the inverted diagonal block into s_u.  Each of the two is
      // loaded whole before its LDS writes, so the step waits for memory twice
      // rather than once per element.
      const HODLR_GLOBAL TLU* inv = DI + int64_t(b) * 2 * RB * RB + (pass == 0 ? 0 : RB * RB);
      constexpr int KL = COMPACT ? HK : RB;   // k rows of the inverse held at once
      constexpr int NQ = RB * CB / kThreads, CS = kThreads / RB, ND = RB * KL / kThreads;
      static_assert(kThreads % RB == 0 && (RB * KL) % kThreads == 0, "fill shapes");
      // (the column index t / RB is one wave's: kept in scalar registers, so
      // each load is a scalar column base plus the thread's row)
      const int fr = threadIdx.x % RB, fc = __builtin_amdgcn_readfirstlane(threadIdx.x / RB);
      MT xv[NQ];
      TLU dv[ND];
      if (pass == 0) {
        const int src = fr < nr ? (perm ? perm[r0 + fr] : r0 + fr) : 0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const int c = fc + q * CS;
          const HODLR_GLOBAL TIN* col = Xin + int64_t(c0 + c) * ldin;
          xv[q] = MT(*(fr < nr && c < nc ? col + src : zero_page<TIN>()));
        }
      } else {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const int c = fc + q * CS;
          const HODLR_GLOBAL TOUT* col = Xout + int64_t(c0 + c) * ldout;
          xv[q] = MT(*(fr < nr && c < nc ? col + (r0 + fr) : zero_page<TOUT>()));
        }
      }
#pragma unroll
      for (int q = 0; q < NQ; ++q) s_t[fr * PT + fc + q * CS] = xv[q];
#pragma unroll
      for (int q = 0; q < ND; ++q) dv[q] = inv[threadIdx.x + q * kThreads];
#pragma unroll
      for (int q = 0; q < ND; ++q) {
        const int e = threadIdx.x + q * kThreads;
        s_u[(e / RB) * P + e % RB] = dv[q];   // As[k][m] = inv(m = e % RB, k = e / RB)
      }
      __syncthreads();
      // s_t -= the product with the solved rows: all reads, then all writes
      {
        typedef typename Mfma<UT>::acc_e AE;
        MT old[TileU::WM * TileU::WN * 4];
        int k = 0;
        tu.for_each([&](int r, int c, AE) { old[k++] = s_t[r * PT + c]; });
        k = 0;
        tu.for_each([&](int r, int c, AE v) { s_t[r * PT + c] = old[k++] - MT(v); });
      }
      __syncthreads();
      t.zero();
      t.template mma<KL>(s_u, s_t, P, PT);
      __syncthreads();   // every wave has read s_u (and s_t)
      if constexpr (COMPACT) {
#pragma unroll
        for (int q = 0; q < ND; ++q) dv[q] = inv[RB * HK + threadIdx.x + q * kThreads];   // columns HK .. RB-1
#pragma unroll
        for (int q = 0; q < ND; ++q) {
          const int e = threadIdx.x + q * kThreads;
          s_u[(e / RB) * P + e % RB] = dv[q];
        }
        __syncthreads();
        t.template mma<HK>(s_u, s_t + HK * PT, P, PT);
        __syncthreads();
      }
      t.for_each([&](int r, int c, typename Mfma<MT>::acc_e v) { s_t[r * PT + c] = v; });
      __syncthreads();
#pragma unroll
      for (int q = 0; q < NQ; ++q) xv[q] = s_t[fr * PT + fc + q * CS];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int c = fc + q * CS;
        if (fr < nr && c < nc) (Xout + int64_t(c0 + c) * ldout)[r0 + fr] = TOUT(xv[q]);
      }
      __syncthreads();
    }
  }
}

// TRSM with 128-row steps: the product with the rows already solved runs as
// one 128 x CB MFMA tile (half the re-reads of solved rows of 64-row steps),
// then the two 64-row halves are finished with the inverted 64 x 64 diagonal
// blocks and the 64 x 64 coupling block between them.  CB = 64 (fp64) or 128
// (fp32) columns per workgroup.
template <class TLU, class TIN, class TOUT, int CB>
__global__ __launch_bounds__(kThreads) void trsm128_kernel(const TrsmDesc* __restrict__ descs, int ndesc,
                                                           const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  typedef TLU MT;
  constexpr int BK = kGemmBigBK, RB = 128, HB = kDB, PT = CB + 1, PI = HB + 1;
  typedef Tile<MT, RB, CB, BK> Big;
  typedef Tile<MT, HB, CB, BK> Half;
  constexpr int SU = (Big::LDS_ELEMS > HB * PI) ? Big::LDS_ELEMS : HB * PI;
  __shared__ MT s_u[SU];       // MFMA stages, then a 64 x 64 block in As layout s_u[k*PI + m]
  __shared__ MT s_t[RB * PT];  // right-hand-side rows of the step, s_t[r*PT + c]

  const int64_t bid = xcd_tile(blockIdx.x, gridDim.x);
  const TrsmDesc d = descs[find_desc(descs, ndesc, bid)];
  const int c0 = int(bid - d.tile_begin) * CB;
  const int m = d.m, ncols = d.ncols;
  const int nc = min(CB, ncols - c0);
  const HODLR_GLOBAL TLU* LU = glob((const TLU*)d.LU);
  const HODLR_GLOBAL TLU* DI = glob((const TLU*)d.dinv);
  const HODLR_GLOBAL TIN* Xin = glob((const TIN*)d.Xin);
  HODLR_GLOBAL TOUT* Xout = glob((TOUT*)d.Xout);
  const HODLR_GLOBAL int32_t* perm = d.perm ? glob(d.perm) : nullptr;
  const int64_t ldlu = d.ldlu, ldin = d.ldin, ldout = d.ldout;
  const int nstep = (m + RB - 1) / RB;

  // s_u <- 64 x 64 block of LU at (r0, k0) (zero outside nr x nk)
  auto load_block = [&](int r0, int k0, int nr, int nk) {
    for (int e = threadIdx.x; e < HB * HB; e += kThreads) {
      const int i = e % HB, k = e / HB;
      s_u[k * PI + i] = (i < nr && k < nk) ? MT(LU[(r0 + i) + int64_t(k0 + k) * ldlu]) : MT(0);
    }
  };
  auto load_inv = [&](int b, int which) {
    const HODLR_GLOBAL TLU* inv = DI + int64_t(b) * 2 * HB * HB + which * HB * HB;
    for (int e = threadIdx.x; e < HB * HB; e += kThreads) s_u[(e / HB) * PI + e % HB] = inv[e];
  };
  // rows [h*64, h*64+64) of s_t <- inv * rows of s_t; also written to Xout
  // (coalesced: 64 rows of a column per wave)
  auto finish_half = [&](int h, int R0, int nr_h) {
    Half t;
    t.zero();
    t.template mma<HB>(s_u, s_t + h * HB * PT, PI, PT);
    __syncthreads();   // every wave has read the half before it is overwritten
    t.for_each([&](int r, int c, typename Mfma<MT>::acc_e v) { s_t[(h * HB + r) * PT + c] = v; });
    __syncthreads();
    for (int e = threadIdx.x; e < HB * CB; e += kThreads) {
      const int r = e % HB, c = e / HB;
      if (r < nr_h && c < nc) Xout[(R0 + h * HB + r) + int64_t(c0 + c) * ldout] = TOUT(s_t[(h * HB + r) * PT + c]);
    }
    __syncthreads();   // s_t is read before the next step refills it
  };
  // rows of half `dst` of s_t -= block(s_u) * rows of half `src`
  auto couple = [&](int dst, int src) {
    Half t;
    t.zero();
    t.template mma<HB>(s_u, s_t + src * HB * PT, PI, PT);
    t.for_each([&](int r, int c, typename Mfma<MT>::acc_e v) { s_t[(dst * HB + r) * PT + c] -= v; });
    __syncthreads();
  };

  for (int pass = 0; pass < 2; ++pass) {      // 0: forward with L, 1: backward with U
    for (int s = 0; s < nstep; ++s) {
      const int st = pass == 0 ? s : nstep - 1 - s;
      const int R0 = st * RB, nr = min(RB, m - R0);
      const int n0 = min(HB, nr), n1 = nr - n0;   // rows of the two halves
      {
        Big t;
        t.zero();
        if (pass == 0)
          gemm_loop<MT, RB, CB, BK, false>(t, s_u, LU, ldlu, R0, m, Xout, ldout, c0, ncols, 0, R0);
        else
          gemm_loop<MT, RB, CB, BK, false>(t, s_u, LU, ldlu, R0, R0 + nr, Xout, ldout, c0, ncols, R0 + RB, m);
        // right-hand-side rows read coalesced into s_t, then s_t -= product
        for (int e = threadIdx.x; e < RB * CB; e += kThreads) {
          const int r = e % RB, c = e / RB;
          MT x = 0;
          if (r < nr && c < nc) {
            if (pass == 0) {
              const int src = perm ? perm[R0 + r] : R0 + r;
              x = MT(Xin[src + int64_t(c0 + c) * ldin]);
            } else {
              x = MT(Xout[(R0 + r) + int64_t(c0 + c) * ldout]);
            }
          }
          s_t[r * PT + c] = x;
        }
        __syncthreads();
        t.for_each([&](int r, int c, typename Mfma<MT>::acc_e v) { s_t[r * PT + c] -= v; });
      }
      const int b0 = 2 * st;   // 64-row diagonal block of the first half
      if (pass == 0) {
        load_inv(b0, 0);
        __syncthreads();
        finish_half(0, R0, n0);
        if (n1 > 0) {
          load_block(R0 + HB, R0, n1, n0);   // L coupling block
          __syncthreads();
          couple(1, 0);
          load_inv(b0 + 1, 0);
          __syncthreads();
          finish_half(1, R0, n1);
        }
      } else {
        if (n1 > 0) {
          load_inv(b0 + 1, 1);
          __syncthreads();
          finish_half(1, R0, n1);
          load_block(R0, R0 + HB, n0, n1);   // U coupling block
          __syncthreads();
          couple(0, 1);
        }
        load_inv(b0, 1);
        __syncthreads();
        finish_half(0, R0, n0);
      }
    }
  }
}

// ============================================================ bf16 LDS TRSM
// bf16 plans: one workgroup per (block, CB-column panel) computes Y = L^-1 P X_in
// and then X_out = U^-1 Y, 64 rows per step, with the panel's solved rows kept
// in LDS as bf16 -- the operand precision of the products with solved rows --
// so a step reads from memory only the factor's block row (from the
// transposed bf16 copy: 16-byte k-contiguous fragments go straight into the
// MFMA A registers, issued PF k-steps ahead) and its own right-hand-side rows.
// The inverted 64 x 64 diagonal block is applied in fp32 (16x16x4 MFMA).
//   s_y: solved rows, octet-planar: element (k, c) at s_y[((k/8)*CB + c)*8 + k%8]
//   s_t: the step's 64 rows, fp32 [r][c]
// Wave w owns rows 16w .. 16w+15 of a step and all CB columns.
__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// bf16x3 product of fp32 operands: a*b ~ ah*bh + ah*bl + al*bh with the hi/lo
// bf16 splits x = xh + xl (relative error ~2^-16, the small terms added first)
__device__ __forceinline__ void split_bf16(const float* x, bf16x8& h, bf16x8& l) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const bf16 xh = bf16(x[e]);
    h[e] = xh;
    l[e] = bf16(x[e] - float(xh));
  }
}

// Waves: 4 row groups of 16 rows x CB/32 column groups of 32 columns, so a
// 64-column panel runs 8 waves (two per SIMD) and the waves of a SIMD hide
// each other's latencies.
template <class TIN, int MMAX, int CB>
__global__ __launch_bounds__(CB * 8) __attribute__((amdgpu_waves_per_eu(2, 2)))
void trsm_lds_kernel(const TrsmDesc* __restrict__ descs, int ndesc, const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  constexpr int NT = CB * 8;
  // PF: factor fragments in flight per wave, enough for a whole step of
  // the largest blocks; with the right-hand side and the inverse fragments
  // the loads in flight stay under the 63 the wait counter can express
  constexpr int RB = kDB, NJ = 2, PT = CB + 1, NQ = CB * RB / NT, ND = 16;
  constexpr int PF = MMAX >= 1024 ? 24 : MMAX / 32;
  static_assert(RB == 64 && (CB == 32 || CB == 64), "16-row x 32-column waves");
  static_assert(PF + NQ + ND <= 63, "loads in flight");
  __shared__ __attribute__((aligned(16))) bf16 s_y[MMAX * CB];
  __shared__ float s_t[RB * PT];
  __shared__ int s_perm[MMAX];   // the row gather of the forward pass, read once

  const int64_t bid = xcd_tile(blockIdx.x, gridDim.x);
  const TrsmDesc d = descs[find_desc(descs, ndesc, bid)];
  const int c0 = int(bid - d.tile_begin) * CB;
  const int m = d.m, nc = min(CB, d.ncols - c0);
  for (int i = threadIdx.x; i < m; i += NT) s_perm[i] = d.perm ? d.perm[i] : i;
  lds_barrier();
  const HODLR_GLOBAL bf16* LT = glob((const bf16*)d.LUt);
  const int64_t ldt = d.ldt;
  const HODLR_GLOBAL float* DI = glob((const float*)d.dinv);
  const int nblk = (m + RB - 1) / RB;
  const int kend = (m + 31) & ~31;
  // (the wave index in a scalar register: every per-wave condition and
  // fragment base below is then scalar)
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lq = lane >> 4;
  const int wr = wave & 3, wc = 32 * (wave >> 2);   // the wave's row group and first column
  // coalesced fill / write-back: thread owns row fr of columns fc + CS*q; the
  // column bases are computed once (a column past nc reads the zero page)
  const int fr = tid % RB, fc = __builtin_amdgcn_readfirstlane(tid / RB);
  constexpr int CS = NT / RB;
  const HODLR_GLOBAL TIN* xin_col[NQ];
  HODLR_GLOBAL float* xout_col[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int c = min(fc + CS * q, nc - 1);
    xin_col[q] = glob((const TIN*)d.Xin) + int64_t(c0 + c) * d.ldin;
    xout_col[q] = glob((float*)d.Xout) + int64_t(c0 + c) * d.ldout;
  }
  const HODLR_GLOBAL bf16x8* zero8 = (const HODLR_GLOBAL bf16x8*)zero_page<double>();
  const int wrow = 16 * wr + li;   // the thread's row of a step in the MFMA A layout

  // per-step register set, reloaded in place for the next step as soon as
  // its values are consumed (no second set, no copies)
  bf16x8 ar[PF];         // the step's factor fragments
  float dv[ND];          // the inverted diagonal block, A fragments (row wrow, k = 32kk + 8lq + e)
  TIN xr[NQ];            // forward right-hand side (X_in rows gathered through perm)
  float xb[NQ];          // backward right-hand side (the forward result)

  auto blk = [&](int t) { return t < nblk ? t : 2 * nblk - 1 - t; };
  // the factor's MFMA A fragment of rows 16g .. 16g+15, k 32ks .. 32ks+31:
  // 1 KB, this lane's 16 bytes at `lane` (the copy is stored fragment by
  // fragment, so every load is one contiguous 1 KB per wave)
  const int nks = int(ldt >> 5);
  auto frag = [&](int g, int ks) {
    return (const HODLR_GLOBAL bf16x8*)LT + (int64_t(g) * nks + ks) * 64 + lane;
  };
  auto load_step = [&](int t) {        // factor fragments of step t's update
    const int r0 = blk(t) * RB;
    const int kb = t < nblk ? 0 : r0 + RB, ke = t < nblk ? r0 : kend;
    const int nk = ke > kb ? (ke - kb) >> 5 : 0;
    const bool ok = r0 + 16 * wr < m;
    const HODLR_GLOBAL bf16x8* p = frag(ok ? (r0 >> 4) + wr : 0, kb >> 5);
    // a row group past the block's last row gets zero fragments: its k loop
    // then adds nothing (registers left unloaded could hold NaN bit
    // patterns, which the zero columns of the inverse would not cancel)
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      if (ok && u < nk) ar[u] = p[64 * u];
      else if (!ok) ar[u] = bf16x8{};
    }
  };
  auto load_dinv = [&](int t) {
    const HODLR_GLOBAL float* inv = DI + int64_t(blk(t)) * 2 * RB * RB + (t < nblk ? 0 : RB * RB) + wrow;
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
      for (int e = 0; e < 8; ++e) dv[kk * 8 + e] = inv[RB * (32 * kk + 8 * lq + e)];
  };
  auto load_rhs_fwd = [&](int t) {
    const int r0 = blk(t) * RB, nr = min(RB, m - r0);
    const int src = s_perm[min(r0 + fr, m - 1)];
#pragma unroll
    for (int q = 0; q < NQ; ++q) xr[q] = *(fr < nr && fc + CS * q < nc ? xin_col[q] + src : zero_page<TIN>());
  };
  auto load_rhs_bwd = [&](int t) {
    const int r0 = blk(t) * RB, nr = min(RB, m - r0);
#pragma unroll
    for (int q = 0; q < NQ; ++q) xb[q] = *(fr < nr && fc + CS * q < nc ? xout_col[q] + (r0 + fr) : zero_page<float>());
  };

  // One step: rows r0 .. r0+63 of the forward (t < nblk) or backward solve.
  auto step = [&](int t, auto fwd) {
    constexpr bool FWD = decltype(fwd)::value;
    const int b = blk(t), r0 = b * RB, nr = min(RB, m - r0);
    const int kb = FWD ? 0 : r0 + RB, ke = FWD ? r0 : kend;
    const int nk = ke > kb ? (ke - kb) >> 5 : 0;
    const bool rowok = r0 + 16 * wr < m;
    const HODLR_GLOBAL bf16x8* arow = frag(rowok ? (r0 >> 4) + wr : 0, kb >> 5);
    // update: acc = factor block row x solved rows (bf16 MFMA, fp32 sums);
    // fragments beyond PF (only in the last steps of the largest blocks) are
    // loaded into the slot just consumed
    f32x4 acc[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[j] = f32x4{0, 0, 0, 0};
    auto kstep = [&](int i, const bf16x8& a) {
      const bf16* yb = s_y + ((((kb >> 3) + 4 * i + lq) * CB) + wc + li) * 8;
      bf16x8 bv[NJ];
#pragma unroll
      for (int j = 0; j < NJ; ++j) bv[j] = *(const bf16x8*)(yb + 16 * j * 8);
#pragma unroll
      for (int j = 0; j < NJ; ++j) acc[j] = mfma_bf16(a, bv[j], acc[j]);
    };
    int i0 = 0;
    for (; i0 + PF <= nk; i0 += PF) {
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        kstep(i0 + u, ar[u]);
        ar[u] = *(rowok && i0 + u + PF < nk ? arow + 64 * (i0 + u + PF) : zero8);
      }
    }
#pragma unroll
    for (int u = 0; u < PF; ++u)
      if (i0 + u < nk) kstep(i0 + u, ar[u]);
    // T = rhs - update (the first backward step's rows are the last forward
    // step's output, still in s_t); the freed registers take the next step's
    // loads, which arrive while this step finishes
    if (FWD) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) s_t[fr * PT + fc + CS * q] = float(xr[q]);
      if (t + 1 < nblk) load_rhs_fwd(t + 1);
    } else {
      if (t != nblk) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) s_t[fr * PT + fc + CS * q] = xb[q];
      }
      if (t + 1 < 2 * nblk) load_rhs_bwd(t + 1);
    }
    if (t + 1 < 2 * nblk) load_step(t + 1);
    lds_barrier();
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) s_t[(16 * wr + 4 * lq + v) * PT + wc + 16 * j + li] -= acc[j][v];
    lds_barrier();
    // X = inv * T, bf16x3
    f32x4 acc2[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc2[j] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      bf16x8 ah, al;
      split_bf16(dv + 8 * kk, ah, al);
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        float y[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) y[e] = s_t[(32 * kk + 8 * lq + e) * PT + wc + 16 * j + li];
        bf16x8 bh, bl;
        split_bf16(y, bh, bl);
        acc2[j] = mfma_bf16(al, bh, acc2[j]);
        acc2[j] = mfma_bf16(ah, bl, acc2[j]);
        acc2[j] = mfma_bf16(ah, bh, acc2[j]);
      }
    }
    if (t + 1 < 2 * nblk) load_dinv(t + 1);
    lds_barrier();
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) s_t[(16 * wr + 4 * lq + v) * PT + wc + 16 * j + li] = acc2[j][v];
    lds_barrier();
    // write back: X_out coalesced, and the rows into s_y as bf16 octets
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      if (fr < nr && fc + CS * q < nc) xout_col[q][r0 + fr] = s_t[fr * PT + fc + CS * q];
    for (int p = tid; p < (RB / 8) * CB; p += NT) {
      const int o = p / CB, c = p % CB;
      bf16x8 v;
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = bf16(s_t[(8 * o + e) * PT + c]);
      *(bf16x8*)(s_y + (((r0 >> 3) + o) * CB + c) * 8) = v;
    }
    lds_barrier();
  };
  load_rhs_fwd(0);
  load_dinv(0);
  load_step(0);
  for (int t = 0; t < nblk; ++t) step(t, std::true_type());
  // the backward steps read rows of X_out the forward steps wrote: one full
  // barrier (global memory included) between the two passes
  __syncthreads();
  for (int t = nblk; t < 2 * nblk; ++t) step(t, std::false_type());
}

// LU factors -> the bf16 copy trsm_lds_kernel reads, stored as MFMA A
// fragments: element (i, k) (row i, column k of LU) at
//   (((i/16) * (ldd/32) + k/32) * 64 + (i%16) + 16*((k%32)/8)) * 8 + k%8
// so the 16 rows x 32 k of one fragment are 1 KB in lane order.  Rows up to
// the next multiple of 16 and columns up to ldd are zero past m.  64 x 64
// tiles through LDS.
__global__ __launch_bounds__(kThreads) void transpose_bf16_kernel(const TransDesc* __restrict__ descs, int ndesc) {
  constexpr int TB = 64;
  __shared__ float s[TB * (TB + 1)];
  const int64_t bid = blockIdx.x;
  const TransDesc d = descs[find_desc(descs, ndesc, bid)];
  const int64_t t = bid - d.tile_begin;
  const int i0 = int(t / d.tiles_k) * TB, k0 = int(t % d.tiles_k) * TB;
  const HODLR_GLOBAL float* src = glob((const float*)d.src);
  HODLR_GLOBAL bf16x8* dst = (HODLR_GLOBAL bf16x8*)glob((bf16*)d.dst);
  const int m = d.m, mr = (m + 15) & ~15;
  const int64_t nks = d.ldd >> 5;
  const int tid = threadIdx.x, r = tid % TB, w = tid / TB;
  float v[TB * TB / kThreads];
#pragma unroll
  for (int q = 0; q < TB * TB / kThreads; ++q) {
    const int k = w + (kThreads / TB) * q;
    v[q] = *(i0 + r < m && k0 + k < m ? src + (i0 + r) + int64_t(k0 + k) * d.lds : zero_page<float>());
  }
#pragma unroll
  for (int q = 0; q < TB * TB / kThreads; ++q) s[(w + (kThreads / TB) * q) * (TB + 1) + r] = v[q];
  __syncthreads();
  // 8 fragments per tile (4 row groups x 2 k steps), one 16-byte octet per
  // thread and pass: fragment f, lane l -> row 16(f%4) + l%16, k 32(f/4) + 8(l/16)
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    const int f = tid / 64 + 4 * pass, l = tid % 64;
    const int i = 16 * (f % 4) + (l % 16), kq = 32 * (f / 4) + 8 * (l / 16);
    if (i0 + i < mr && k0 + kq < d.ldd) {
      bf16x8 o;
#pragma unroll
      for (int e = 0; e < 8; ++e) o[e] = bf16(s[(kq + e) * (TB + 1) + i]);
      dst[((int64_t((i0 + i) >> 4) * nks + ((k0 + kq) >> 5)) * 64) + l] = o;
    }
  }
}

// ====================================================================== norms
__global__ __launch_bounds__(kThreads) void norm_kernel(const NormDesc* __restrict__ descs, int ndesc,
                                                        const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  const int64_t bid = blockIdx.x;
  const NormDesc d = descs[find_desc(descs, ndesc, bid)];
  const int64_t chunk = bid - d.tile_begin;
  const int64_t b = chunk * kNormChunk, e = min(d.n, b + kNormChunk);
  double s = 0.0;
  const HODLR_GLOBAL double* x = glob(d.x);
  for (int64_t i = b + threadIdx.x; i < e; i += kThreads) s += x[i] * x[i];
  __shared__ double red[kThreads];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) d.partial[chunk] = red[0];
}

__global__ void norm_finish_kernel(const NormDesc* __restrict__ descs, int ndesc, const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ndesc) return;
  const NormDesc d = descs[i];
  double s = 0.0;
  for (int q = 0; q < d.nchunks; ++q) s += d.partial[q];
  *d.out = sqrt(s);
}

// Convergence check of the refinement loop, on the device, so the sweeps of a
// captured solve need no host round trip: relres = the largest ||r_j|| / ||b_j||
// over every right-hand side of every system (||r_j|| where ||b_j|| = 0, +inf
// for NaN).  mode 0 follows the first residual (iter = 0, relres_first set);
// mode 1 ends a sweep (iter += 1).  Then active = !(relres <= tol) && iter <
// maxiter: the gate of the next sweep.  One workgroup.
__global__ __launch_bounds__(kThreads) void refine_check_kernel(const CheckDesc* __restrict__ descs, int ndesc,
                                                                RefineCtl* ctl, int mode,
                                                                const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  double worst = 0.0;
  for (int s = threadIdx.x; s < ndesc; s += kThreads) {
    const CheckDesc d = descs[s];
    for (int j = 0; j < d.c; ++j) {
      const double rn = d.rnorm[j], bn = d.bnorm[j];
      double v = bn > 0.0 ? rn / bn : rn;
      if (!(v == v)) v = INFINITY;
      worst = fmax(worst, v);
    }
  }
  __shared__ double red[kThreads];
  red[threadIdx.x] = worst;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double rr = red[0];
    const int iter = mode == 0 ? 0 : ctl->iter + 1;
    ctl->iter = iter;
    ctl->relres = rr;
    if (mode == 0) ctl->relres_first = rr;
    ctl->active = (!(rr <= ctl->tol) && iter < ctl->maxiter) ? 1 : 0;
  }
}

// ================================================================== launchers
namespace {
inline unsigned grid1(int64_t tiles) { return unsigned(tiles); }
}

hipError_t launch_gemm(DT ta, DT tb, DT tc, DT mt, bool trans, int role, int tile, const GemmDesc* d, int n,
                       int64_t tiles, hipStream_t s, const int32_t* gate) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  dim3 g(grid1(tiles)), b(kThreads);
#define HODLR_GEMM(R, A, B, C, M, T, BM, BN) \
  hipLaunchKernelGGL((gemm_kernel<R, A, B, C, M, T, BM, BN>), g, b, 0, s, d, n, gate)
  // one of the three tilings: 0 = 64 x 64, 1 = 128 x 128, 2 = 64 x 32
#define HODLR_GEMM3(R, A, B, C, M, T)                                  \
  do {                                                                 \
    if (tile == 1) HODLR_GEMM(R, A, B, C, M, T, 128, 128);             \
    else if (tile == 2) HODLR_GEMM(R, A, B, C, M, T, 64, kGemmNarrowBN); \
    else HODLR_GEMM(R, A, B, C, M, T, 64, 64);                         \
  } while (0)
  if (ta == F64 && tb == F64 && tc == F64 && mt == F64) {
    if (role == 4) {
      if (tile == 2) {
        if (trans) HODLR_GEMM(4, double, double, double, double, true, 64, kGemmNarrowBN);
        else HODLR_GEMM(4, double, double, double, double, false, 64, kGemmNarrowBN);
      } else {
        if (trans) HODLR_GEMM(4, double, double, double, double, true, 64, 64);
        else HODLR_GEMM(4, double, double, double, double, false, 64, 64);
      }
    } else if (trans) HODLR_GEMM3(1, double, double, double, double, true);
    else if (role == 2) HODLR_GEMM3(2, double, double, double, double, false);
    else if (role == 3) HODLR_GEMM3(3, double, double, double, double, false);
    else if (tile == 2) HODLR_GEMM(0, double, double, double, double, false, 64, kGemmNarrowBN);
    else HODLR_GEMM(0, double, double, double, double, false, 64, 64);
  } else if (ta == F64 && tb == F32 && tc == F64 && mt == F32 && trans) {
    HODLR_GEMM3(1, double, float, double, float, true);
  } else if (ta == F32 && tb == F64 && tc == F32 && mt == F32 && !trans) {
    HODLR_GEMM3(2, float, double, float, float, false);
  } else if (ta == F32 && tb == F32 && tc == F32 && mt == F32 && !trans) {
    if (role == 2) HODLR_GEMM3(2, float, float, float, float, false);
    else if (role == 3) HODLR_GEMM3(3, float, float, float, float, false);
    else if (tile == 2) HODLR_GEMM(0, float, float, float, float, false, 64, kGemmNarrowBN);
    else HODLR_GEMM(0, float, float, float, float, false, 64, 64);
  } else if (ta == F32 && tb == F32 && tc == F64 && mt == F32 && trans) {
    HODLR_GEMM3(1, float, float, double, float, true);
  } else if (ta == F32 && tb == F32 && tc == F64 && mt == BF16 && trans) {
    HODLR_GEMM3(1, float, float, double, bf16, true);
  } else if (ta == F64 && tb == F32 && tc == F64 && mt == BF16 && trans) {
    HODLR_GEMM3(1, double, float, double, bf16, true);
  } else if (ta == F32 && tb == F32 && tc == F32 && mt == BF16 && !trans) {
    if (role == 2) HODLR_GEMM3(2, float, float, float, bf16, false);
    else HODLR_GEMM3(3, float, float, float, bf16, false);
  } else {
    return hipErrorInvalidValue;
  }
#undef HODLR_GEMM3
#undef HODLR_GEMM
  return hipGetLastError();
}

hipError_t launch_panel_gemm(DT ts, DT mt, bool bfrag, const GemmDesc* d, int n, int64_t tiles, hipStream_t s,
                             const int32_t* gate) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  if (ts != F32 || mt != BF16) return hipErrorInvalidValue;
  auto k = bfrag ? panel_gemm_kernel<true> : panel_gemm_kernel<false>;
  hipLaunchKernelGGL(k, dim3(grid1(tiles)), dim3(kThreads), 0, s, d, n, gate);
  return hipGetLastError();
}

hipError_t launch_stream(int ncols_max, bool vec, const StreamDesc* d, int n, int64_t tiles, hipStream_t s,
                         const int32_t* gate) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  if (ncols_max > kStreamCols) return hipErrorInvalidValue;
  dim3 g(grid1(tiles)), b(kThreads);
  const bool two = ncols_max > 16;
  auto k = two ? (vec ? stream_kernel<2, true> : stream_kernel<2, false>)
               : (vec ? stream_kernel<1, true> : stream_kernel<1, false>);
  hipLaunchKernelGGL(k, g, b, 0, s, d, n, gate);
  return hipGetLastError();
}

hipError_t launch_frag_bf16(bool vec, const FragDesc* d, int n, int64_t tiles, hipStream_t s, const int32_t* gate) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  auto k = vec ? frag_bf16_kernel<true> : frag_bf16_kernel<false>;
  hipLaunchKernelGGL(k, dim3(grid1(tiles)), dim3(kThreads), 0, s, d, n, gate);
  return hipGetLastError();
}

hipError_t launch_reduce_bf16(const GemmDesc* d, int n, int64_t tiles, hipStream_t s, const int32_t* gate) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_bf16_kernel, dim3(grid1(tiles)), dim3(kThreads), 0, s, d, n, gate);
  return hipGetLastError();
}

hipError_t launch_sum(const SumDesc* d, int n, int64_t tiles, hipStream_t s, const int32_t* gate) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(sum_kernel, dim3(grid1(tiles)), dim3(kThreads), 0, s, d, n, gate);
  return hipGetLastError();
}

hipError_t launch_copy(DT ts, DT td, const CopyDesc* d, int n, int64_t tiles, hipStream_t s, const int32_t* gate) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  dim3 g(grid1(tiles)), b(kThreads);
  if (ts == F64 && td == F64) hipLaunchKernelGGL((copy_kernel<double, double>), g, b, 0, s, d, n, gate);
  else if (ts == F64 && td == F32) hipLaunchKernelGGL((copy_kernel<double, float>), g, b, 0, s, d, n, gate);
  else if (ts == F32 && td == F64) hipLaunchKernelGGL((copy_kernel<float, double>), g, b, 0, s, d, n, gate);
  else if (ts == F32 && td == F32) hipLaunchKernelGGL((copy_kernel<float, float>), g, b, 0, s, d, n, gate);
  else if (ts == F64 && td == BF16) hipLaunchKernelGGL((copy_kernel<double, bf16>), g, b, 0, s, d, n, gate);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_lu_panel(DT t, const LuDesc* d, int n, int k0, int rows_max, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (rows_max > kPanelThreads) return hipErrorInvalidValue;
  dim3 g(n);
#define HODLR_PANEL(NT, RPT)                                                                         \
  do {                                                                                               \
    if (t == F64) hipLaunchKernelGGL((lu_panel_kernel<double, NT, RPT>), g, dim3(NT), 0, s, d, k0); \
    else if (n > kPanelOccBlocks)                                                                    \
      hipLaunchKernelGGL((lu_panel_kernel<float, NT, RPT, PanelOcc<float, NT, RPT>::value>), g,      \
                         dim3(NT), 0, s, d, k0);                                                     \
    else hipLaunchKernelGGL((lu_panel_kernel<float, NT, RPT>), g, dim3(NT), 0, s, d, k0);           \
  } while (0)
  if (rows_max <= 4 * 64) {
    // short panels: one wave per block (no cross-wave pivot combine, and four
    // times as many blocks in flight per CU)
    if (rows_max <= 64) HODLR_PANEL(64, 1);
    else if (rows_max <= 128) HODLR_PANEL(64, 2);
    else HODLR_PANEL(64, 4);
  } else if (rows_max <= kPanelNT) HODLR_PANEL(kPanelNT, 1);
  else if (rows_max <= 2 * kPanelNT) HODLR_PANEL(kPanelNT, 2);
  else HODLR_PANEL(kPanelNT, 4);
#undef HODLR_PANEL
  return hipGetLastError();
}

hipError_t launch_lu_dinv(DT t, const LuDesc* d, int n, int b0, int nblk, hipStream_t s) {
  if (n <= 0 || nblk <= 0) return hipSuccess;
  dim3 g(nblk, n), b(128);
  if (t == F64) hipLaunchKernelGGL((lu_dinv_kernel<double>), g, b, 0, s, d, b0);
  else hipLaunchKernelGGL((lu_dinv_kernel<float>), g, b, 0, s, d, b0);
  return hipGetLastError();
}

hipError_t launch_lu_small(DT t, int mmax, const LuDesc* d, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (mmax > kSmallLU) return hipErrorInvalidValue;
  if (mmax <= 64) {
    // one wave per block: 8 x 8 threads, 8 x 8 entries each
    if (t == F64) hipLaunchKernelGGL((lu_small_kernel<double, 8, 8>), dim3(n), dim3(64), 0, s, d);
    else hipLaunchKernelGGL((lu_small_kernel<float, 8, 8>), dim3(n), dim3(64), 0, s, d);
  } else {
    constexpr int G = kSmallLUGrid, NU = kSmallLU / kSmallLUGrid;
    if (t == F64) hipLaunchKernelGGL((lu_small_kernel<double, G, NU>), dim3(n), dim3(G * G), 0, s, d);
    else hipLaunchKernelGGL((lu_small_kernel<float, G, NU>), dim3(n), dim3(G * G), 0, s, d);
  }
  return hipGetLastError();
}

hipError_t launch_lu_perm(const LuDesc* d, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(lu_perm_kernel, dim3(n), dim3(64), 0, s, d);
  return hipGetLastError();
}

hipError_t launch_lu_init(const LuDesc* d, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(lu_init_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d, n);
  return hipGetLastError();
}

hipError_t launch_trsm(DT tlu, DT tin, DT tout, DT ut, int variant, const TrsmDesc* d, int n, int64_t tiles,
                       hipStream_t s, const int32_t* gate) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  dim3 g(grid1(tiles)), b(kThreads);
  if (ut == BF16) {
    if (variant != 0 || tlu != F32 || tout != F32) return hipErrorInvalidValue;
    if (tin == F64) hipLaunchKernelGGL((trsm_kernel<float, double, float, bf16>), g, b, 0, s, d, n, gate);
    else if (tin == F32) hipLaunchKernelGGL((trsm_kernel<float, float, float, bf16>), g, b, 0, s, d, n, gate);
    else return hipErrorInvalidValue;
  } else if (variant == 0) {
    if (tlu == F64 && tin == F64 && tout == F64)
      hipLaunchKernelGGL((trsm_kernel<double, double, double, double>), g, b, 0, s, d, n, gate);
    else if (tlu == F32 && tin == F64 && tout == F32)
      hipLaunchKernelGGL((trsm_kernel<float, double, float, float>), g, b, 0, s, d, n, gate);
    else if (tlu == F32 && tin == F32 && tout == F32)
      hipLaunchKernelGGL((trsm_kernel<float, float, float, float>), g, b, 0, s, d, n, gate);
    else return hipErrorInvalidValue;
  } else {
    if (tlu == F64 && tin == F64 && tout == F64)
      hipLaunchKernelGGL((trsm128_kernel<double, double, double, 64>), g, b, 0, s, d, n, gate);
    else if (tlu == F32 && tin == F64 && tout == F32)
      hipLaunchKernelGGL((trsm128_kernel<float, double, float, 128>), g, b, 0, s, d, n, gate);
    else if (tlu == F32 && tin == F32 && tout == F32)
      hipLaunchKernelGGL((trsm128_kernel<float, float, float, 128>), g, b, 0, s, d, n, gate);
    else return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_trsm_lds(DT tin, int variant, int mmax, const TrsmDesc* d, int n, int64_t tiles, hipStream_t s,
                           const int32_t* gate) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  if (mmax > 1024 || (variant != 2 && variant != 3) || (tin != F64 && tin != F32)) return hipErrorInvalidValue;
  dim3 g(grid1(tiles)), b(kThreads);
#define HODLR_TRSM_LDS(TI, MM, CB) hipLaunchKernelGGL((trsm_lds_kernel<TI, MM, CB>), g, dim3(CB * 8), 0, s, d, n, gate)
#define HODLR_TRSM_LDS2(MM, CB)                      \
  do {                                              \
    if (tin == F64) HODLR_TRSM_LDS(double, MM, CB); \
    else HODLR_TRSM_LDS(float, MM, CB);             \
  } while (0)
  if (mmax <= 256) {
    if (variant == 2) HODLR_TRSM_LDS2(256, 64);
    else HODLR_TRSM_LDS2(256, 32);
  } else {
    if (variant == 2) HODLR_TRSM_LDS2(1024, 64);
    else HODLR_TRSM_LDS2(1024, 32);
  }
#undef HODLR_TRSM_LDS2
#undef HODLR_TRSM_LDS
  return hipGetLastError();
}

hipError_t launch_transpose_bf16(const TransDesc* d, int n, int64_t tiles, hipStream_t s) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(transpose_bf16_kernel, dim3(grid1(tiles)), dim3(kThreads), 0, s, d, n);
  return hipGetLastError();
}

hipError_t launch_norm(const NormDesc* d, int n, int64_t tiles, hipStream_t s, const int32_t* gate) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(norm_kernel, dim3(grid1(tiles)), dim3(kThreads), 0, s, d, n, gate);
  return hipGetLastError();
}

hipError_t launch_norm_finish(const NormDesc* d, int n, hipStream_t s, const int32_t* gate) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(norm_finish_kernel, dim3((n + 63) / 64), dim3(64), 0, s, d, n, gate);
  return hipGetLastError();
}

hipError_t launch_refine_check(const CheckDesc* d, int n, RefineCtl* ctl, int mode, hipStream_t s,
                               const int32_t* gate) {
  if (n <= 0 || !ctl) return hipErrorInvalidValue;
  hipLaunchKernelGGL(refine_check_kernel, dim3(1), dim3(kThreads), 0, s, d, n, ctl, mode, gate);
  return hipGetLastError();
}

}  // namespace hodlr

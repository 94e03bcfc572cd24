// Batched gfx950 kernels of the HODLR factor-and-solve path.
//
//   gemm_kernel      C = alpha*op(A)*B + beta*C over a descriptor batch, MFMA
//                    16x16x4 (fp64 or fp32), optional deterministic split-K.
//                    Serves the level reduce V^T [d|u] (src/solver/gemm.cc:315-384),
//                    the broadcast d -= u*eta (gemm.cc:417-478), the node-solve
//                    products (src/solver/solver_tasks.cc:236-302), the LU
//                    trailing updates and the HODLR matvec.
//   lu_*_kernel      batched blocked LU with partial pivoting (getrf semantics)
//                    of the leaf blocks (solver_tasks.cc:437) and of the r x r
//                    Schur complements (solver_tasks.cc:281).
//   trsm_kernel      X_out = U^-1 L^-1 P X_in per (block, 64-column panel),
//                    left-looking, row interchanges applied while gathering X_in.
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.hpp"
#include "kernels.hpp"

namespace hodlr {

// ======================================================================= GEMM
// ROLE only separates the symbol names of the callers in profiles:
// 0 node/matvec products, 1 reduce V^T X, 2 broadcast d -= u eta, 3 LU trailing update.
template <int ROLE, class TA, class TB, class TC, class MT, bool TRANS, int BM, int BN>
__global__ __launch_bounds__(kThreads) void gemm_kernel(const GemmDesc* __restrict__ descs, int ndesc) {
  constexpr int BK = kGemmBK;
  typedef Tile<MT, BM, BN, BK> TileT;
  __shared__ MT lds[TileT::LDS_ELEMS];
  MT* As = lds;
  MT* Bs = lds + BK * TileT::PA;

  const int64_t bid = blockIdx.x;
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
  const TA* A = (const TA*)d.A;
  const TB* B = (const TB*)d.B;

  TileT t;
  t.zero();
  for (int k0 = kb; k0 < ke; k0 += BK) {
    load_a<MT, BM, BK, TRANS>(As, A, d.lda, r0, d.m, k0, ke);
    load_b<MT, BN, BK>(Bs, B, d.ldb, k0, ke, c0, d.n);
    __syncthreads();
    t.mma(As, Bs);
    __syncthreads();
  }
  const double alpha = d.alpha, beta = d.beta;
  const int m = d.m, n = d.n;
  if (d.nsplit > 1) {
    TC* out = (TC*)d.Cout + int64_t(split) * d.split_stride;
    const int64_t ldo = d.ldcout;
    t.for_each([&](int r, int c, typename Mfma<MT>::acc_e v) {
      int gr = r0 + r, gc = c0 + c;
      if (gr < m && gc < n) out[gr + gc * ldo] = TC(alpha * double(v));
    });
  } else {
    TC* out = (TC*)d.Cout;
    const TC* cin = (const TC*)d.Cin;
    const int64_t ldo = d.ldcout, ldi = d.ldcin;
    t.for_each([&](int r, int c, typename Mfma<MT>::acc_e v) {
      int gr = r0 + r, gc = c0 + c;
      if (gr < m && gc < n) {
        double x = alpha * double(v);
        if (beta != 0.0) x += beta * double(cin[gr + gc * ldi]);
        out[gr + gc * ldo] = TC(x);
      }
    });
  }
}

// split-K chunks are added in chunk order: deterministic, no atomics
__global__ __launch_bounds__(kThreads) void sum_kernel(const SumDesc* __restrict__ descs, int ndesc) {
  const int64_t bid = blockIdx.x;
  const SumDesc d = descs[find_desc(descs, ndesc, bid)];
  const int64_t base = (bid - d.tile_begin) * kSumTile;
  const int64_t total = int64_t(d.m) * d.n;
  const double* P = (const double*)d.P;
  double* out = (double*)d.out;
  for (int e = threadIdx.x; e < kSumTile; e += kThreads) {
    int64_t idx = base + e;
    if (idx >= total) break;
    double s = 0.0;
    for (int q = 0; q < d.nsplit; ++q) s += P[q * d.stride + idx];
    int64_t i = idx % d.m, j = idx / d.m;
    out[i + j * d.ldo] = s;
  }
}

// ======================================================================= copy
template <class TS, class TD>
__global__ __launch_bounds__(kThreads) void copy_kernel(const CopyDesc* __restrict__ descs, int ndesc) {
  const int64_t bid = blockIdx.x;
  const CopyDesc d = descs[find_desc(descs, ndesc, bid)];
  const int64_t base = (bid - d.tile_begin) * kCopyTile;
  const int64_t total = int64_t(d.rows) * d.cols;
  const TS* src = (const TS*)d.src;
  TD* dst = (TD*)d.dst;
  for (int e = threadIdx.x; e < kCopyTile; e += kThreads) {
    int64_t idx = base + e;
    if (idx >= total) break;
    int64_t i = idx % d.rows, j = idx / d.rows;
    double x = d.alpha * double(src[i + j * d.lds]);
    if (d.beta != 0.0) x += d.beta * double(dst[i + j * d.ldd]);
    if (i == j) x += d.diag;
    dst[i + j * d.ldd] = TD(x);
  }
}

// ========================================================================= LU
// Panel of kLuNB columns starting at k0, factored as two 32-column
// sub-panels held in registers (thread t owns rows t, t+256, ... of the
// sub-panel).  Partial pivoting picks the first largest |a| (idamax); the
// column below a zero pivot is left unscaled and info records it, as dgetf2.
constexpr int kPW = 32;

template <class T>
__device__ __forceinline__ void pivot_combine(T& bv, int& bi, T ov, int oi) {
  if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
}

template <class T, int RPT>
__global__ __launch_bounds__(kThreads) void lu_panel_kernel(const LuDesc* __restrict__ descs, int k0) {
  const LuDesc d = descs[blockIdx.x];
  const int m = d.m;
  if (k0 >= m) return;
  T* A = (T*)d.A;
  const int64_t lda = d.lda;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nb = min(kLuNB, m - k0);

  __shared__ T s_rowp[2][kPW];
  __shared__ T s_rowj[2][kPW];
  __shared__ T s_rv[2][4];
  __shared__ int s_ri[2][4];
  __shared__ int s_piv[kLuNB];        // absolute pivot row of panel column jg
  __shared__ T s_l[kPW * (kPW + 1)];  // L11 of sub-panel 0
  __shared__ T s_u[kPW * (kPW + 1)];  // U12 block
  __shared__ int s_info;
  if (tid == 0) s_info = 0;
  __syncthreads();

  for (int sp = 0; sp < 2; ++sp) {
    const int ks = k0 + kPW * sp;
    if (ks >= k0 + nb) break;
    const int w = min(kPW, k0 + nb - ks);
    const int rows = m - ks;
    T a[RPT][kPW];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int q = tid + kThreads * i;
#pragma unroll
      for (int jj = 0; jj < kPW; ++jj)
        a[i][jj] = (q < rows && jj < w) ? A[(ks + q) + int64_t(ks + jj) * lda] : T(0);
    }
#pragma unroll
    for (int j = 0; j < kPW; ++j) {
      if (j < w) {
        const int buf = j & 1;
        T bv = T(-1);
        int bi = 0x7fffffff;
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
          const int q = tid + kThreads * i;
          if (q >= j && q < rows) {
            T v = fabs(a[i][j]);
            if (v > bv) { bv = v; bi = q; }
          }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          T ov = __shfl_xor(bv, off);
          int oi = __shfl_xor(bi, off);
          pivot_combine(bv, bi, ov, oi);
        }
        if (lane == 0) { s_rv[buf][wave] = bv; s_ri[buf][wave] = bi; }
        __syncthreads();
        T pv = s_rv[buf][0];
        int p = s_ri[buf][0];
#pragma unroll
        for (int q = 1; q < 4; ++q) pivot_combine(pv, p, s_rv[buf][q], s_ri[buf][q]);
        // a column of NaNs offers no candidate: keep row j (no interchange)
        if (p >= rows) p = j;
        // publish the pivot row and row j
        const int ip = p / kThreads;
        if (tid == p % kThreads) {
#pragma unroll
          for (int i = 0; i < RPT; ++i)
            if (i == ip) {
#pragma unroll
              for (int jj = 0; jj < kPW; ++jj) s_rowp[buf][jj] = a[i][jj];
            }
        }
        if (tid == j) {
#pragma unroll
          for (int jj = 0; jj < kPW; ++jj) s_rowj[buf][jj] = a[0][jj];
        }
        if (tid == 0) {
          s_piv[kPW * sp + j] = ks + p;
          d.ipiv[ks + j] = ks + p;
          if (!(pv > T(0)) && s_info == 0) s_info = ks + j + 1;
        }
        __syncthreads();
        if (p != j) {
          if (tid == j) {
#pragma unroll
            for (int jj = 0; jj < kPW; ++jj) a[0][jj] = s_rowp[buf][jj];
          }
          if (tid == p % kThreads) {
#pragma unroll
            for (int i = 0; i < RPT; ++i)
              if (i == ip) {
#pragma unroll
                for (int jj = 0; jj < kPW; ++jj) a[i][jj] = s_rowj[buf][jj];
              }
          }
        }
        const T piv = s_rowp[buf][j];
        const T rinv = piv != T(0) ? T(1) / piv : T(0);
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
          const int q = tid + kThreads * i;
          if (q > j && q < rows) {
            const T l = piv != T(0) ? a[i][j] * rinv : a[i][j];
            a[i][j] = l;
#pragma unroll
            for (int jj = j + 1; jj < kPW; ++jj) a[i][jj] -= l * s_rowp[buf][jj];
          }
        }
      }
    }
    // write the factored sub-panel back
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int q = tid + kThreads * i;
      if (q < rows) {
#pragma unroll
        for (int jj = 0; jj < kPW; ++jj)
          if (jj < w) A[(ks + q) + int64_t(ks + jj) * lda] = a[i][jj];
      }
    }
    __syncthreads();
    if (sp == 0 && nb > kPW) {
      const int nc2 = nb - kPW;   // columns of sub-panel 1
      // sub-panel 0 interchanges on sub-panel 1 columns
      if (tid < nc2) {
        const int64_t col = int64_t(ks + kPW + tid) * lda;
        for (int j = 0; j < w; ++j) {
          const int pr = s_piv[j];
          if (pr != ks + j) {
            T x = A[(ks + j) + col];
            A[(ks + j) + col] = A[pr + col];
            A[pr + col] = x;
          }
        }
      }
      if (tid < kPW) {
#pragma unroll
        for (int jj = 0; jj < kPW; ++jj) s_l[tid * (kPW + 1) + jj] = a[0][jj];
      }
      __syncthreads();
      for (int e = tid; e < kPW * nc2; e += kThreads) {
        const int r = e % kPW, c = e / kPW;
        s_u[r * (kPW + 1) + c] = A[(ks + r) + int64_t(ks + kPW + c) * lda];
      }
      __syncthreads();
      // U12 = L11^-1 A12 (unit lower), one thread per column
      if (tid < nc2) {
        for (int i = 1; i < kPW; ++i) {
          T x = s_u[i * (kPW + 1) + tid];
          for (int k = 0; k < i; ++k) x -= s_l[i * (kPW + 1) + k] * s_u[k * (kPW + 1) + tid];
          s_u[i * (kPW + 1) + tid] = x;
        }
      }
      __syncthreads();
      for (int e = tid; e < kPW * nc2; e += kThreads) {
        const int r = e % kPW, c = e / kPW;
        A[(ks + r) + int64_t(ks + kPW + c) * lda] = s_u[r * (kPW + 1) + c];
      }
      // A22 -= L21 U12 on the sub-panel 1 columns, L21 rows from registers
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const int q = tid + kThreads * i;
        if (q >= kPW && q < rows) {
          for (int c = 0; c < nc2; ++c) {
            T x = A[(ks + q) + int64_t(ks + kPW + c) * lda];
#pragma unroll
            for (int k = 0; k < kPW; ++k) x -= a[i][k] * s_u[k * (kPW + 1) + c];
            A[(ks + q) + int64_t(ks + kPW + c) * lda] = x;
          }
        }
      }
      __syncthreads();
    }
    if (sp == 1) {
      // sub-panel 1 interchanges on sub-panel 0 columns
      if (tid < kPW) {
        const int64_t col = int64_t(k0 + tid) * lda;
        for (int j = 0; j < w; ++j) {
          const int pr = s_piv[kPW + j];
          if (pr != ks + j) {
            T x = A[(ks + j) + col];
            A[(ks + j) + col] = A[pr + col];
            A[pr + col] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  // panel interchanges on every column outside the panel (dlaswp)
  const int nout = m - nb;
  for (int c = tid; c < nout; c += kThreads) {
    const int col = c < k0 ? c : c + nb;
    const int64_t cb = int64_t(col) * lda;
    for (int j = 0; j < nb; ++j) {
      const int pr = s_piv[j];
      if (pr != k0 + j) {
        T x = A[(k0 + j) + cb];
        A[(k0 + j) + cb] = A[pr + cb];
        A[pr + cb] = x;
      }
    }
  }
  if (tid == 0 && s_info != 0 && *d.info == 0) *d.info = s_info;
}

// U12 = L11^-1 A12 for the kLuNB rows of the panel at k0, 64 columns per block.
template <class T>
__global__ __launch_bounds__(kThreads) void lu_rowblock_kernel(const LuDesc* __restrict__ descs, int k0) {
  const LuDesc d = descs[blockIdx.y];
  const int m = d.m;
  if (k0 >= m) return;
  const int nb = min(kLuNB, m - k0);
  const int c0 = k0 + nb + blockIdx.x * 64;
  if (c0 >= m) return;
  const int nc = min(64, m - c0);
  T* A = (T*)d.A;
  const int64_t lda = d.lda;
  constexpr int P = 65;
  __shared__ T s_l[kLuNB * P];
  __shared__ T s_x[kLuNB * P];
  for (int e = threadIdx.x; e < nb * nb; e += kThreads) {
    const int r = e % nb, k = e / nb;
    s_l[r * P + k] = A[(k0 + r) + int64_t(k0 + k) * lda];
  }
  for (int e = threadIdx.x; e < nb * nc; e += kThreads) {
    const int r = e % nb, c = e / nb;
    s_x[r * P + c] = A[(k0 + r) + int64_t(c0 + c) * lda];
  }
  __syncthreads();
  if (threadIdx.x < nc) {
    const int c = threadIdx.x;
    for (int i = 1; i < nb; ++i) {
      T x = s_x[i * P + c];
      for (int k = 0; k < i; ++k) x -= s_l[i * P + k] * s_x[k * P + c];
      s_x[i * P + c] = x;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nb * nc; e += kThreads) {
    const int r = e % nb, c = e / nb;
    A[(k0 + r) + int64_t(c0 + c) * lda] = s_x[r * P + c];
  }
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

// ======================================================================= TRSM
// One workgroup per (block, 64-column panel): forward elimination with L
// (unit lower) on P*X_in, then back substitution with U.  Each 64-row step
// first applies the already solved rows with an MFMA product, then solves the
// 64 x 64 diagonal block in LDS, one thread per column.
template <class TLU, class TIN, class TOUT>
__global__ __launch_bounds__(kThreads) void trsm_kernel(const TrsmDesc* __restrict__ descs, int ndesc) {
  typedef TLU MT;
  constexpr int BK = kGemmBK, RB = kTrsmRB, CB = kTrsmCB, P = RB + 1;
  typedef Tile<MT, RB, CB, BK> TileT;
  constexpr int UNION = (TileT::LDS_ELEMS > RB * P) ? TileT::LDS_ELEMS : RB * P;
  __shared__ MT s_u[UNION];   // GEMM staging, then the diagonal block
  __shared__ MT s_t[RB * P];  // right-hand-side block, s_t[r*P + c]
  MT* As = s_u;
  MT* Bs = s_u + BK * TileT::PA;

  const int64_t bid = blockIdx.x;
  const TrsmDesc d = descs[find_desc(descs, ndesc, bid)];
  const int c0 = int(bid - d.tile_begin) * CB;
  const int m = d.m, ncols = d.ncols;
  const int nc = min(CB, ncols - c0);
  const TLU* LU = (const TLU*)d.LU;
  const TIN* Xin = (const TIN*)d.Xin;
  TOUT* Xout = (TOUT*)d.Xout;
  const int64_t ldlu = d.ldlu, ldin = d.ldin, ldout = d.ldout;
  const int nblk = (m + RB - 1) / RB;
  TileT t;

  // ---- forward: Y = L^-1 P X_in
  for (int b = 0; b < nblk; ++b) {
    const int r0 = b * RB, nr = min(RB, m - r0);
    t.zero();
    for (int k0 = 0; k0 < r0; k0 += BK) {
      load_a<MT, RB, BK, false>(As, LU, ldlu, r0, m, k0, r0);
      load_b<MT, CB, BK>(Bs, Xout, ldout, k0, r0, c0, ncols);
      __syncthreads();
      t.mma(As, Bs);
      __syncthreads();
    }
    t.for_each([&](int r, int c, typename Mfma<MT>::acc_e v) {
      MT x = 0;
      if (r < nr && c < nc) {
        const int src = d.perm ? d.perm[r0 + r] : r0 + r;
        x = MT(Xin[src + int64_t(c0 + c) * ldin]) - v;
      }
      s_t[r * P + c] = x;
    });
    for (int e = threadIdx.x; e < RB * RB; e += kThreads) {
      const int r = e % RB, k = e / RB;
      s_u[r * P + k] = (r < nr && k < nr) ? MT(LU[(r0 + r) + int64_t(r0 + k) * ldlu]) : MT(0);
    }
    __syncthreads();
    if (threadIdx.x < nc) {
      const int c = threadIdx.x;
      for (int i = 1; i < nr; ++i) {
        MT x = s_t[i * P + c];
        for (int k = 0; k < i; ++k) x -= s_u[i * P + k] * s_t[k * P + c];
        s_t[i * P + c] = x;
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < RB * nc; e += kThreads) {
      const int r = e % RB, c = e / RB;
      if (r < nr) Xout[(r0 + r) + int64_t(c0 + c) * ldout] = TOUT(s_t[r * P + c]);
    }
    __syncthreads();
  }
  // ---- backward: X = U^-1 Y
  for (int b = nblk - 1; b >= 0; --b) {
    const int r0 = b * RB, nr = min(RB, m - r0);
    t.zero();
    for (int k0 = r0 + RB; k0 < m; k0 += BK) {
      load_a<MT, RB, BK, false>(As, LU, ldlu, r0, r0 + nr, k0, m);
      load_b<MT, CB, BK>(Bs, Xout, ldout, k0, m, c0, ncols);
      __syncthreads();
      t.mma(As, Bs);
      __syncthreads();
    }
    t.for_each([&](int r, int c, typename Mfma<MT>::acc_e v) {
      MT x = 0;
      if (r < nr && c < nc) x = MT(Xout[(r0 + r) + int64_t(c0 + c) * ldout]) - v;
      s_t[r * P + c] = x;
    });
    for (int e = threadIdx.x; e < RB * RB; e += kThreads) {
      const int r = e % RB, k = e / RB;
      s_u[r * P + k] = (r < nr && k < nr) ? MT(LU[(r0 + r) + int64_t(r0 + k) * ldlu]) : MT(0);
    }
    __syncthreads();
    if (threadIdx.x < nc) {
      const int c = threadIdx.x;
      for (int i = nr - 1; i >= 0; --i) {
        MT x = s_t[i * P + c];
        for (int k = i + 1; k < nr; ++k) x -= s_u[i * P + k] * s_t[k * P + c];
        s_t[i * P + c] = x / s_u[i * P + i];
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < RB * nc; e += kThreads) {
      const int r = e % RB, c = e / RB;
      if (r < nr) Xout[(r0 + r) + int64_t(c0 + c) * ldout] = TOUT(s_t[r * P + c]);
    }
    __syncthreads();
  }
}

// ====================================================================== norms
__global__ __launch_bounds__(kThreads) void norm_kernel(const NormDesc* __restrict__ descs, int ndesc) {
  const int64_t bid = blockIdx.x;
  const NormDesc d = descs[find_desc(descs, ndesc, bid)];
  const int64_t chunk = bid - d.tile_begin;
  const int64_t b = chunk * kNormChunk, e = min(d.n, b + kNormChunk);
  double s = 0.0;
  for (int64_t i = b + threadIdx.x; i < e; i += kThreads) s += d.x[i] * d.x[i];
  __shared__ double red[kThreads];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) d.partial[chunk] = red[0];
}

__global__ void norm_finish_kernel(const NormDesc* __restrict__ descs, int ndesc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ndesc) return;
  const NormDesc d = descs[i];
  double s = 0.0;
  for (int q = 0; q < d.nchunks; ++q) s += d.partial[q];
  *d.out = sqrt(s);
}

// ================================================================== launchers
namespace {
inline unsigned grid1(int64_t tiles) { return unsigned(tiles); }
}

hipError_t launch_gemm(DT ta, DT tb, DT tc, DT mt, bool trans, int role, const GemmDesc* d, int n, int64_t tiles,
                       hipStream_t s) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  dim3 g(grid1(tiles)), b(kThreads);
  constexpr int BM = kGemmBM, BN = kGemmBN;
  if (ta == F64 && tb == F64 && tc == F64 && mt == F64) {
    if (trans) hipLaunchKernelGGL((gemm_kernel<1, double, double, double, double, true, BM, BN>), g, b, 0, s, d, n);
    else if (role == 2) hipLaunchKernelGGL((gemm_kernel<2, double, double, double, double, false, BM, BN>), g, b, 0, s, d, n);
    else if (role == 3) hipLaunchKernelGGL((gemm_kernel<3, double, double, double, double, false, BM, BN>), g, b, 0, s, d, n);
    else hipLaunchKernelGGL((gemm_kernel<0, double, double, double, double, false, BM, BN>), g, b, 0, s, d, n);
  } else if (ta == F64 && tb == F32 && tc == F64 && mt == F32 && trans) {
    hipLaunchKernelGGL((gemm_kernel<1, double, float, double, float, true, BM, BN>), g, b, 0, s, d, n);
  } else if (ta == F32 && tb == F64 && tc == F32 && mt == F32 && !trans) {
    hipLaunchKernelGGL((gemm_kernel<2, float, double, float, float, false, BM, BN>), g, b, 0, s, d, n);
  } else if (ta == F32 && tb == F32 && tc == F32 && mt == F32 && !trans) {
    hipLaunchKernelGGL((gemm_kernel<3, float, float, float, float, false, BM, BN>), g, b, 0, s, d, n);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_sum(const SumDesc* d, int n, int64_t tiles, hipStream_t s) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(sum_kernel, dim3(grid1(tiles)), dim3(kThreads), 0, s, d, n);
  return hipGetLastError();
}

hipError_t launch_copy(DT ts, DT td, const CopyDesc* d, int n, int64_t tiles, hipStream_t s) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  dim3 g(grid1(tiles)), b(kThreads);
  if (ts == F64 && td == F64) hipLaunchKernelGGL((copy_kernel<double, double>), g, b, 0, s, d, n);
  else if (ts == F64 && td == F32) hipLaunchKernelGGL((copy_kernel<double, float>), g, b, 0, s, d, n);
  else if (ts == F32 && td == F64) hipLaunchKernelGGL((copy_kernel<float, double>), g, b, 0, s, d, n);
  else if (ts == F32 && td == F32) hipLaunchKernelGGL((copy_kernel<float, float>), g, b, 0, s, d, n);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_lu_panel(DT t, const LuDesc* d, int n, int k0, int rows_max, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int rpt = (rows_max + kThreads - 1) / kThreads;
  dim3 g(n), b(kThreads);
  if (rpt > 4) return hipErrorInvalidValue;
  if (t == F64) {
    if (rpt <= 1) hipLaunchKernelGGL((lu_panel_kernel<double, 1>), g, b, 0, s, d, k0);
    else if (rpt == 2) hipLaunchKernelGGL((lu_panel_kernel<double, 2>), g, b, 0, s, d, k0);
    else hipLaunchKernelGGL((lu_panel_kernel<double, 4>), g, b, 0, s, d, k0);
  } else {
    if (rpt <= 1) hipLaunchKernelGGL((lu_panel_kernel<float, 1>), g, b, 0, s, d, k0);
    else if (rpt == 2) hipLaunchKernelGGL((lu_panel_kernel<float, 2>), g, b, 0, s, d, k0);
    else hipLaunchKernelGGL((lu_panel_kernel<float, 4>), g, b, 0, s, d, k0);
  }
  return hipGetLastError();
}

hipError_t launch_lu_rowblock(DT t, const LuDesc* d, int n, int k0, int mmax, hipStream_t s) {
  const int ncols = mmax - k0 - kLuNB;
  if (n <= 0 || ncols <= 0) return hipSuccess;
  dim3 g((ncols + 63) / 64, n), b(kThreads);
  if (t == F64) hipLaunchKernelGGL((lu_rowblock_kernel<double>), g, b, 0, s, d, k0);
  else hipLaunchKernelGGL((lu_rowblock_kernel<float>), g, b, 0, s, d, k0);
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

hipError_t launch_trsm(DT tlu, DT tin, DT tout, const TrsmDesc* d, int n, int64_t tiles, hipStream_t s) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  dim3 g(grid1(tiles)), b(kThreads);
  if (tlu == F64 && tin == F64 && tout == F64) hipLaunchKernelGGL((trsm_kernel<double, double, double>), g, b, 0, s, d, n);
  else if (tlu == F32 && tin == F64 && tout == F32) hipLaunchKernelGGL((trsm_kernel<float, double, float>), g, b, 0, s, d, n);
  else if (tlu == F32 && tin == F32 && tout == F32) hipLaunchKernelGGL((trsm_kernel<float, float, float>), g, b, 0, s, d, n);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_norm(const NormDesc* d, int n, int64_t tiles, hipStream_t s) {
  if (n <= 0 || tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(norm_kernel, dim3(grid1(tiles)), dim3(kThreads), 0, s, d, n);
  return hipGetLastError();
}

hipError_t launch_norm_finish(const NormDesc* d, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(norm_finish_kernel, dim3((n + 63) / 64), dim3(64), 0, s, d, n);
  return hipGetLastError();
}

}  // namespace hodlr

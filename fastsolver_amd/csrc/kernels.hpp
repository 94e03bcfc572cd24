// Host-side launchers of the batched gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "desc.hpp"

namespace hodlr {

// BF16: the math type of the bf16 plans' products; as a storage type only for the
// bf16 copies of V (copy destination, reduce operand)
enum DT { F64 = 0, F32 = 1, BF16 = 2 };

constexpr int kGemmBM = 64, kGemmBN = 64, kGemmBK = 16;
constexpr int kGemmBK32 = 16;       // k step of the 64-wide tiles in fp32
constexpr int kGemmBigBK = 16;      // k step of the 128 x 128 tiles and of the 128-row TRSM
constexpr int kGemmBKbf16 = 32;     // k step of bf16 tiles (one 16x16x32 MFMA)
constexpr int kTrsmBK = 16, kTrsmBK32 = 16;   // k step of the 64-row TRSM (fp64, fp32)
constexpr int kGemmNarrowBN = 32;   // tile 2: 64 x 32 for problems at most 32 columns wide
constexpr int kGemmNarrowBK = 32;   // k step of the narrow tiles (HBM-bound streaming products)
constexpr int kTrsmRB = 64, kTrsmCB = 64;
constexpr int kLuNB = 64;          // panel width of the blocked LU
constexpr int kLuStep = 512;       // columns per trailing update (a power-of-two multiple of kLuNB)
constexpr int kCopyTile = 1024;    // elements per copy workgroup
constexpr int kSumTile = 1024;
constexpr int64_t kNormChunk = 65536;

// Supported (A, B, C, compute) type tuples:
//   (F64,F64,F64,F64) N and T;  (F64,F32,F64,F32) T;  (F32,F64,F32,F32) N;  (F32,F32,F32,F32) N;
//   (F64,F32,F64,BF16) T;  (F32,F32,F32,BF16) N.
// tile 0: 64 x 64 workgroup tiles, 1: 128 x 128 (roles 1-3 only)
// `gate` (every launcher that takes one): a device flag, the launch does nothing while it reads 0
hipError_t launch_gemm(DT ta, DT tb, DT tc, DT mt, bool trans_a, int role, int tile, const GemmDesc* d, int n,
                       int64_t tiles, hipStream_t s, const int32_t* gate = nullptr);
// row-panel GEMM (k <= 128, fp32 storage, bf16 MFMA): one workgroup per 64-row panel of C
// sweeping all columns; B's leading dimension a multiple of 8 with zero rows past k, or with
// `bfrag` B as bf16 MFMA fragments in the FragDesc layout of B (ldb = its column-group count);
// descriptors tiled by 64-row panels
constexpr int kPanelKMax = 128;
hipError_t launch_panel_gemm(DT ts, DT mt, bool bfrag, const GemmDesc* d, int n, int64_t tiles, hipStream_t s,
                             const int32_t* gate = nullptr);
// fp64 -> bf16 MFMA A fragments of the transpose (FragDesc), 4 kFragPerWave fragments per
// workgroup (one column group, consecutive row blocks); `vec`: 32-byte loads, allowed when every
// src is 32-byte aligned, every lds a multiple of 4 and every rows a multiple of 8
constexpr int kFragPerWave = 4;
hipError_t launch_frag_bf16(bool vec, const FragDesc* d, int n, int64_t tiles, hipStream_t s,
                            const int32_t* gate = nullptr);
// bf16 plans' reduce W = A^T B (GemmDesc, trans): A in the FragDesc fragment layout of A^T
// (lda = its column-group count), B fp32, C fp64; tiles of kReduceRows x kReduceCols, split-K
// chunks (kchunk) a multiple of 32 rows
constexpr int kReduceRows = 128, kReduceCols = 32;
hipError_t launch_reduce_bf16(const GemmDesc* d, int n, int64_t tiles, hipStream_t s, const int32_t* gate = nullptr);
// FP64 streaming products (StreamDesc: C = alpha (A1 B1 + A2 B2) + beta Cin), tiles of kStreamRows
// rows x stream_tile_cols(ncols_max) columns; `ncols_max` = the widest product of the batch (at most
// kStreamCols); `vec`: 32-byte vector loads, allowed when every A and B pointer is 32-byte aligned,
// every leading dimension and every m a multiple of 4
constexpr int kStreamRows = 256, kStreamCols = 32;
inline int stream_tile_cols(int ncols_max) { return ncols_max > 16 ? 32 : 16; }
hipError_t launch_stream(int ncols_max, bool vec, const StreamDesc* d, int n, int64_t tiles, hipStream_t s,
                         const int32_t* gate = nullptr);
hipError_t launch_sum(const SumDesc* d, int n, int64_t tiles, hipStream_t s,
                      const int32_t* gate = nullptr);  // double partials
hipError_t launch_copy(DT ts, DT td, const CopyDesc* d, int n, int64_t tiles, hipStream_t s,
                       const int32_t* gate = nullptr);
hipError_t launch_lu_panel(DT t, const LuDesc* d, int n, int k0, int rows_max, hipStream_t s);
hipError_t launch_lu_perm(const LuDesc* d, int n, hipStream_t s);
// whole LU (in place, ipiv, info; not dinv / perm) of blocks of at most kSmallLU rows, in registers
constexpr int kSmallLU = 128;
// thread grid side of the register-resident LU for 65-128 rows: 16 x 16
// threads, 8 x 8 entries each (32 x 32 threads, 16 waves per block, measured
// slower: more waves per SIMD repeat the per-column pivot search and
// exchange); blocks of at most 64 rows take one wave (8 x 8 threads)
constexpr int kSmallLUGrid = 16;
hipError_t launch_lu_small(DT t, int mmax, const LuDesc* d, int n, hipStream_t s);
// inverted 64 x 64 diagonal blocks b0 .. b0+nblk-1 of every factor
hipError_t launch_lu_dinv(DT t, const LuDesc* d, int n, int b0, int nblk, hipStream_t s);
hipError_t launch_lu_init(const LuDesc* d, int n, hipStream_t s);   // info = 0
// (TLU, TIN, TOUT) in {(F64,F64,F64), (F32,F64,F32), (F32,F32,F32)}; compute type = TLU, except the
// products with the solved rows, which run in `ut` (= TLU, or BF16 with TLU = F32, variant 0 only)
// variant 0: 64-row steps, 64-column panels; 1: 128-row steps, trsm_panel_cols(tlu)-column panels
hipError_t launch_trsm(DT tlu, DT tin, DT tout, DT ut, int variant, const TrsmDesc* d, int n, int64_t tiles,
                       hipStream_t s, const int32_t* gate = nullptr);
// variants 2 / 3 (bf16 plans, TLU = F32, ut = BF16): the panel's solved rows kept in LDS as bf16,
// 64 / 32 columns per workgroup, factor rows from the transposed bf16 copy (TrsmDesc::LUt);
// `mmax` is the largest block of the batch (<= 1024)
inline int trsm_panel_cols(DT tlu, int variant) {
  return variant == 0 ? kTrsmCB : variant == 1 ? (tlu == F64 ? 64 : 128) : variant == 2 ? 64 : 32;
}
hipError_t launch_trsm_lds(DT tin, int variant, int mmax, const TrsmDesc* d, int n, int64_t tiles, hipStream_t s,
                           const int32_t* gate = nullptr);
hipError_t launch_transpose_bf16(const TransDesc* d, int n, int64_t tiles, hipStream_t s);
hipError_t launch_norm(const NormDesc* d, int n, int64_t tiles, hipStream_t s, const int32_t* gate = nullptr);
hipError_t launch_norm_finish(const NormDesc* d, int n, hipStream_t s, const int32_t* gate = nullptr);
// refinement convergence check (one workgroup): mode 0 after the first residual, 1 after a sweep
hipError_t launch_refine_check(const CheckDesc* d, int n, RefineCtl* ctl, int mode, hipStream_t s,
                               const int32_t* gate = nullptr);

}  // namespace hodlr

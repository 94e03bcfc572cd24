// Batch descriptors shared by the host schedule and the device kernels.  Every
// batched launch reads an array of descriptors; each descriptor owns a
// contiguous range of flattened workgroup tiles starting at tile_begin.
#pragma once
#include <stdint.h>

namespace hodlr {

// Cout = alpha * op(A) * B + beta * Cin   (op(A) = A or A^T), C is m x n.
// With nsplit > 1 the K range is cut into chunks of kchunk; chunk s writes its
// raw product alpha*op(A)*B to Cout + s*split_stride (leading dimension
// ldcout) and a SumDesc pass adds the chunks in a fixed order.
struct GemmDesc {
  const void* A;
  const void* B;
  const void* Cin;
  void* Cout;
  int64_t lda, ldb, ldcin, ldcout;
  int32_t m, n, k, kchunk;
  int32_t tiles_m, tiles_n, nsplit, pad0;
  int64_t split_stride;
  double alpha, beta;
  int64_t tile_begin;
};

// out = sum_{s < nsplit} P[s*stride + i + j*m]   (fixed order), m x n
// out = sum over q < nsplit of the m x n partial at P + q*stride (leading
// dimension ldp), added in q order.
struct SumDesc {
  const void* P;
  int64_t stride;
  void* out;
  int64_t ldo;
  int32_t m, n, nsplit, ldp;
  int64_t tile_begin;
};

// Batched LU with partial pivoting of square blocks (getrf semantics).
struct LuDesc {
  void* A;            // m x m, column-major, leading dimension lda
  int64_t lda;
  int32_t m, pad0;
  int32_t* ipiv;      // m entries, 0-based row swapped with row i at step i
  int32_t* perm;      // m entries, row i of P*A is row perm[i] of A
  int32_t* info;      // first zero pivot (1-based), 0 if none
  void* dinv;         // per 64-row diagonal block b: L_bb^-1 then U_bb^-1, 64 x 64 each (ld 64)
  int64_t pad1;       // keeps the descriptor at 64 bytes
};

// X_out = U^-1 L^-1 P X_in for one LU-factored block and ncols right-hand sides.
struct TrsmDesc {
  const void* LU;
  int64_t ldlu;
  const void* dinv;     // inverted diagonal blocks of LU (LuDesc::dinv)
  const int32_t* perm;  // may be null: no row interchanges
  const void* Xin;
  int64_t ldin;
  void* Xout;
  int64_t ldout;
  int32_t m, ncols;
  const void* LUt;      // bf16 plans: LU rounded to bf16 in MFMA A-fragment order (TransDesc),
  int64_t ldt;          //   columns padded with zeros to ldt = m rounded up to 32
  int64_t tile_begin;
};

// dst = bf16(src) (m x m, ld lds) in MFMA A-fragment order: 16 rows x 32
// columns per 1 KB fragment, fragments row group by row group, ldd/32 per
// group; zero past m (rows up to a multiple of 16, columns up to ldd).
// One workgroup per 64 x 64 tile, tiles_k tiles along k per row band.
struct TransDesc {
  const void* src;
  int64_t lds;
  void* dst;
  int64_t ldd;
  int32_t m, tiles_k;
  int64_t tile_begin;
};

// dst = alpha * src + beta * dst + diag * I (rows x cols), types given by the
// launch.  beta = 0 never reads dst.
struct CopyDesc {
  const void* src;
  int64_t lds;
  void* dst;
  int64_t ldd;
  int32_t rows, cols;
  double alpha, beta, diag;
  int64_t tile_begin;
};

// Sum of squares of one column segment: partial[chunk] for rows
// [chunk*kNormChunk, ...) of x (n rows).
struct NormDesc {
  const double* x;
  int64_t n;
  double* partial;     // nchunks entries
  double* out;         // out = sqrt(sum of partials), written by the finish pass
  int32_t nchunks, pad0;
  int64_t tile_begin;
};

// FP64 product with few columns (the residual and matvec sweeps):
//   C (m x n) = alpha * (op(A1) B1 + A2 B2) + beta * Cin
// op(A1) = A1 (m x k1) or A1^T (A1: k1 x m); A2 (m x k2, never transposed)
// may be absent (k2 = 0).  The streaming kernel takes op(A1) = A1, one
// workgroup per 256 rows x 16 or 32 columns of C; transposed products go to
// the tile engine.
struct StreamDesc {
  const double* A1;
  const double* B1;
  const double* A2;
  const double* B2;
  const double* Cin;
  double* Cout;
  int64_t lda1, ldb1, lda2, ldb2, ldcin, ldcout;
  int32_t m, n, k1, k2;
  int32_t tiles_m, tiles_n;
  double alpha, beta;
  int64_t tile_begin;
};

// dst = the bf16 MFMA A fragments of src^T (src: rows x cols fp64, ld lds):
// fragment (kb, g) -- rows 32kb .. 32kb+31, columns 16g .. 16g+15 of src --
// is 1 KB at dst + ((kb * ngroups + g) * 64 + lane) * 16 bytes, lane p + 16q
// holding src(32kb + 8q + e, 16g + p) for e = 0..7; zero past rows / cols.
// One workgroup per 4 kFragPerWave fragments of one column group (consecutive kb).
struct FragDesc {
  const double* src;
  int64_t lds;
  void* dst;
  int32_t rows, cols, ngroups, pad0;
  int64_t tile_begin;
};

// Device-side state of the refinement loop of one plan (hodlr_solve).  The
// host writes tol and maxiter before a solve; the check kernel keeps the rest.
struct RefineCtl {
  int32_t active;        // 1: the next sweep runs (the gate of its launches)
  int32_t iter;          // sweeps done
  int32_t maxiter;
  int32_t pad0;
  double tol;
  double relres;         // after the last residual
  double relres_first;   // after the first residual
};

// Residual and right-hand-side 2-norms of one system's c columns.
struct CheckDesc {
  const double* rnorm;
  const double* bnorm;
  int32_t c, pad0;
};

}  // namespace hodlr

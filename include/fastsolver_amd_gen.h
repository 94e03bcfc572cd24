/* Host-side input generation for HODLR systems (libfastsolver_amd_gen.so).
 *
 * Not on the solve path.  It writes the caller's host arrays in the layout
 * the solver consumes:
 *   X  N x (c + R) column-major, ld >= N: [b | U(1) | ... | U(L)]
 *      (the reference's "U region", src/htree/hodlr_matrix.cc:340-352)
 *   V  N x R       column-major, ld >= N: [V(1) | ... | V(L)]
 *   K  leaf diagonal blocks, HODLR_K_PACKED (leaf after leaf, each m x m with
 *      ld m) or HODLR_K_STACKED (leaf block at row row_beg of an N x max_m
 *      array with ld ldk, as the reference's K region, hodlr_matrix.cc:431-445)
 * where R = r_1 + ... + r_L and a node at depth d owns columns
 * c + r_1 + ... + r_{d-1} .. + r_d of X.
 */
#ifndef FASTSOLVER_AMD_GEN_H
#define FASTSOLVER_AMD_GEN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HODLR_K_PACKED 0
#define HODLR_K_STACKED 1

#define HODLR_UV_UNIFORM 0 /* U, V entries uniform[-1,1] / sqrt(rows of the block) */
#define HODLR_UV_DECAY 1   /* U = Q_U diag(10^(-k/decay)), V = Q_V, Q from thin QR of randn */

typedef struct {
  uint64_t seed;      /* Philox key */
  uint32_t system_id; /* independent sub-stream per system of a batch */
  double d_scale;     /* D = randn * d_scale + d_shift * I; d_scale <= 0 means 1/sqrt(m) */
  double d_shift;
  int uv_mode;        /* HODLR_UV_UNIFORM or HODLR_UV_DECAY */
  double decay;       /* decay denominator of HODLR_UV_DECAY */
} hodlr_family;

/* Philox4x32-10 block function (Salmon, Moraes, Dror, Shaw, SC'11). */
void hodlr_gen_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Raw stream values (uniform [0,1) or Box-Muller normal) of one (system, kind,
 * node) stream, elements first .. first+count-1. */
int hodlr_gen_stream_values(uint64_t seed, uint32_t system, uint32_t kind, uint32_t node, int normal,
                            int64_t first, int64_t count, double* out);

/* Which LAPACK the decaying family's QR uses; -1 if none was found. */
int hodlr_gen_lapack_name(char* buf, int len);

/* The reference's test matrix: A = diag*I + U U^T, U(i,j) = (i+j) mod rank, with
 * the drand48 right-hand side reseeded with `seed` at the start of every region
 * (the leaves when rhs_region_rows <= 0, else blocks of rhs_region_rows rows). */
int hodlr_gen_circulant(int64_t n, int64_t threshold, int nrhs, int rank, double diag, int64_t seed,
                        int64_t rhs_region_rows, double* X, int64_t ldx, double* V, int64_t ldv, double* K,
                        int64_t ldk, int klayout);

/* The reference's right-hand side alone (HodlrMatrix::init_rhs, hodlr_matrix.cc:175-209):
 * columns 0 .. nrhs-1 of X, drand48 reseeded with `seed` at the start of every
 * region (the leaves when rhs_region_rows <= 0), drawn row by row. */
int hodlr_gen_rhs_drand48(int64_t n, int64_t threshold, int nrhs, int rank, int64_t seed,
                          int64_t rhs_region_rows, double* X, int64_t ldx);

/* Seeded random family (right-hand side uniform [0,1)). */
int hodlr_gen_random(int64_t n, int64_t threshold, int nrhs, int nranks, const int* ranks,
                     const hodlr_family* fam, double* X, int64_t ldx, double* V, int64_t ldv, double* K,
                     int64_t ldk, int klayout);

/* Tree shape: number of leaves, depth, columns of X and V, doubles of packed K. */
int hodlr_gen_tree(int64_t n, int64_t threshold, int nrhs, int nranks, const int* ranks, int* nleaves,
                   int* maxdepth, int64_t* xcols, int64_t* vcols, int64_t* kpacked);

/* Leaf row ranges and depths in row order (arrays of length nleaves). */
int hodlr_gen_leaves(int64_t n, int64_t threshold, int nrhs, int nranks, const int* ranks, int64_t* row_beg,
                     int64_t* nrow, int* depth);

#ifdef __cplusplus
}
#endif
#endif

/* fastsolver_amd: HODLR factor-and-solve on one MI355X (gfx950).
 *
 * C-ABI of libfastsolver_amd.so.  All array arguments are device pointers to
 * column-major data with explicit leading dimensions; sizes are int64; every
 * entry point returns 0 or a negative HODLR_ERR_* code and never aborts.
 * Work is enqueued on the plan's stream; results are ready when the stream is.
 *
 * Data layout of one HODLR system (the reference's, src/htree/hodlr_matrix.cc):
 *   tree   split while nrow > threshold: children nrow/2 and nrow - nrow/2
 *          (hodlr_matrix.cc:138-169); a node at depth d >= 1 has rank r_d.
 *   X      N x (c + R), [b | U(1) | ... | U(L)]: the right-hand side followed by
 *          the U factor of every depth (the reference's "U region",
 *          hodlr_matrix.cc:340-352); R = r_1 + ... + r_L.
 *   V      N x R, the V factor of every depth (hodlr_matrix.cc:388-472).
 *          Off-diagonal blocks of a node: A01 = U0 V1^T, A10 = U1 V0^T.
 *   K      leaf diagonal blocks, HODLR_K_PACKED (leaf after leaf in row order,
 *          each m x m with ld m) or HODLR_K_STACKED (leaf at row row_beg of an
 *          N x max_m array with ld ldk, hodlr_matrix.cc:431-445).
 * Unlike the reference the solve is not destructive: X, V and K are only read;
 * the factor [x | A_sub^-1 U] lives in the plan's workspace.
 */
#ifndef FASTSOLVER_AMD_H
#define FASTSOLVER_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HODLR_K_PACKED 0
#define HODLR_K_STACKED 1

#define HODLR_PREC_FP64 0 /* every phase in FP64 */
#define HODLR_PREC_FP32 1 /* leaf LU, TRSM, reduce and broadcast in FP32 (MFMA f32), node LU and
                             refinement in FP64 */
#define HODLR_PREC_BF16 2 /* as FP32, with the leaf LU trailing updates, the TRSM products with
                             solved rows, the reduce and the broadcast on BF16 MFMA (operands
                             rounded to BF16, FP32 accumulation and storage) */

#define HODLR_OK 0
#define HODLR_ERR_ARG -1         /* invalid argument */
#define HODLR_ERR_TREE -2        /* tree cannot be built (sizes, ranks, leaf <= rank) */
#define HODLR_ERR_UNSUPPORTED -3 /* leaf or rank larger than the kernels support (1024) */
#define HODLR_ERR_HIP -4         /* a HIP runtime call failed */
#define HODLR_ERR_NOMEM -5       /* device allocation failed */
#define HODLR_ERR_SINGULAR -6    /* a zero pivot was met (see hodlr_status) */
#define HODLR_ERR_NOTCONVERGED -7
#define HODLR_ERR_SCRATCH -8     /* input too large for the fixed scratch buffer (hodlr_dgesv_batched_noalloc) */

typedef struct hodlr_plan_s* hodlr_plan_t;

typedef struct {
  int64_t n;             /* rows N */
  int64_t threshold;     /* leaf size bound (the reference's `threshold`) */
  int32_t nrhs;          /* c, right-hand-side columns */
  int32_t nranks;        /* 1: same rank at every depth; else one rank per depth 1..L */
  const int32_t* ranks;  /* host array of nranks ranks */
  const double* X;       /* device, N x (c + R), ld ldx */
  int64_t ldx;
  const double* V;       /* device, N x R, ld ldv */
  int64_t ldv;
  const double* K;       /* device, leaf blocks */
  int64_t ldk;           /* HODLR_K_STACKED only */
  int32_t klayout;
  int32_t pad0;
  double* x;             /* device, N x c solution, ld ldxs */
  int64_t ldxs;
} hodlr_system_t;

typedef struct {
  int32_t nsys;
  int32_t precision;
  int32_t height;          /* levels above the leaves (max over systems) */
  int32_t nleaves;         /* all systems */
  int64_t workspace_bytes; /* device memory the plan owns */
  double flops_factor;     /* algorithmic flops of one fused factor+solve pass */
  double flops_leaf, flops_reduce, flops_node, flops_bcast;
  double bytes_leaf, bytes_reduce, bytes_bcast; /* algorithmic HBM bytes (element size of the plan) */
  double flops_matvec, bytes_matvec;            /* one FP64 residual */
  double flops_solve_only, bytes_solve_only;    /* one correction solve */
  int32_t launches_factor;                      /* kernel launches of the fused pass */
  int32_t pad0;
} hodlr_plan_info_t;

typedef struct {
  int32_t iterations;      /* refinement sweeps after the first solve */
  int32_t converged;
  double relres;           /* max over systems and columns of ||b - A x|| / ||b|| */
  double relres_first;     /* the same after the first (low-precision) solve */
} hodlr_solve_report_t;

/* Plan: builds every tree, allocates the workspace and uploads the launch
 * descriptors of all phases.  Replaces HodlrMatrix::create_tree and the task
 * registration (hodlr_matrix.cc:44-82, fast_solver.cc:36-50).  `stream` is a
 * hipStream_t (NULL = the default stream). */
int hodlr_plan_create(hodlr_plan_t* plan, const hodlr_system_t* systems, int32_t nsys, int32_t precision,
                      void* stream);
int hodlr_plan_destroy(hodlr_plan_t plan);
int hodlr_plan_query(hodlr_plan_t plan, hodlr_plan_info_t* info);

/* Fused factor + solve of every system, in the plan's precision: leaf solves,
 * then for every height reduce, node solve and broadcast; then x = the first
 * c workspace columns.  Replaces FastSolver::bfs_solve (fast_solver.cc:56-72,
 * 136-255). */
int hodlr_factor_solve(hodlr_plan_t plan);

/* The phases of hodlr_factor_solve, for callers that drive the sweep
 * themselves (heights 1 .. info.height, bottom up):
 *   leaf      solve_legion_leaf on all leaves     (solver_tasks.cc:554-609, 424-438)
 *   reduce    4 x gemm_reduce per node of height h (fast_solver.cc:228-235; gemm.cc:208-229)
 *   node      solve_node_matrix per node           (solver_tasks.cc:57-106, 139-305)
 *   broadcast 2 x gemm_broadcast per node          (fast_solver.cc:250-253; gemm.cc:232-284)
 *   extract   x <- workspace[:, 0:c]                                                      */
int hodlr_leaf_solve(hodlr_plan_t plan);
int hodlr_level_reduce(hodlr_plan_t plan, int32_t height);
int hodlr_level_node_solve(hodlr_plan_t plan, int32_t height);
int hodlr_level_broadcast(hodlr_plan_t plan, int32_t height);
int hodlr_extract_solution(hodlr_plan_t plan);

/* Apply the stored factorization to new right-hand sides (N x c per system):
 * out[s] = A_s^-1 rhs[s].  Needs a preceding hodlr_factor_solve. */
int hodlr_solve_only(hodlr_plan_t plan, const double* const* rhs, const int64_t* ldr, double* const* out,
                     const int64_t* ldo);

/* y[s] = A_s x[s] in FP64 (N x c per system). */
int hodlr_matvec(hodlr_plan_t plan, const double* const* x, const int64_t* ldx, double* const* y,
                 const int64_t* ldy);

/* Relative residual of the current solution: relres[s*c_max + j] =
 * ||b_j - A x_j|| / ||b_j|| (host array of nsys * max_c doubles), and its max. */
int hodlr_residual(hodlr_plan_t plan, double* relres, double* relres_max);

/* Full solve: factor + solve in the plan's precision, then FP64 iterative
 * refinement until max relative residual <= tol or maxiter sweeps.  The
 * convergence check runs on the device and gates the sweeps; with
 * hodlr_use_graph on, the pass, the first residual and as many sweeps as the
 * previous call took are one graph replay (one host synchronisation per
 * solve when the count repeats). */
int hodlr_solve(hodlr_plan_t plan, double tol, int32_t maxiter, hodlr_solve_report_t* report);

/* First zero pivot (1-based, 0 = none) of every leaf block (host array of
 * info.nleaves) and of every node Schur complement (host array of the
 * number of interior nodes, may be NULL).  Synchronises the stream. */
int hodlr_status(hodlr_plan_t plan, int32_t* leaf_info, int32_t* node_info, int32_t* n_nodes);

/* Standalone batched operators: the BLAS/LAPACK calls of the reference's
 * task bodies on raw device pointers (host arrays of `batch` entries), for a
 * caller that keeps its own task structure.  Each call runs on `stream`
 * (a hipStream_t) and synchronises it.
 *   hodlr_dgemm_batched  C = alpha op(A) B + beta C, op = 'n' or 't', FP64:
 *                        dgemm_ of GEMM_Reduce_Task (gemm.cc:376, 't','n'),
 *                        GEMM_Broadcast_Task (gemm.cc:473) and LUSolveTask
 *                        (solver_tasks.cc:236,261,299)
 *   hodlr_sgemm_batched  the same in FP32, op(A) = A
 *   hodlr_dgesv_batched  A X = B with partial pivoting, n <= 1024: dgesv_ of
 *                        serial_leaf_solve (solver_tasks.cc:437) and LUSolveTask
 *                        (solver_tasks.cc:281).  A is overwritten by its LU
 *                        factors, B by X; info[i] as LAPACK (B is not
 *                        meaningful where info[i] > 0). */
int hodlr_dgemm_batched(int32_t batch, const char* transa, const int32_t* m, const int32_t* n, const int32_t* k,
                        const double* alpha, const double* const* A, const int64_t* lda, const double* const* B,
                        const int64_t* ldb, const double* beta, double* const* C, const int64_t* ldc, void* stream);
int hodlr_sgemm_batched(int32_t batch, const int32_t* m, const int32_t* n, const int32_t* k, const double* alpha,
                        const float* const* A, const int64_t* lda, const float* const* B, const int64_t* ldb,
                        const double* beta, float* const* C, const int64_t* ldc, void* stream);
/* hodlr_sgemm_batched with the products on BF16 MFMA (v_mfma_f32_16x16x32_bf16):
 * A and B rounded to BF16, FP32 accumulation and output.  No counterpart in the
 * reference: the BF16 precision of the factorization (HODLR_PREC_BF16). */
int hodlr_sgemm_batched_bf16(int32_t batch, const int32_t* m, const int32_t* n, const int32_t* k,
                             const double* alpha, const float* const* A, const int64_t* lda,
                             const float* const* B, const int64_t* ldb, const double* beta, float* const* C,
                             const int64_t* ldc, void* stream);
int hodlr_dgesv_batched(int32_t batch, const int32_t* n, const int32_t* nrhs, double* const* A, const int64_t* lda,
                        double* const* B, const int64_t* ldb, int32_t* info, void* stream);

/* hodlr_dgesv_batched without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * pivots, permutations, inverted diagonal blocks, copy of B, info words and
 * launch descriptors live in a fixed 64 MiB device array of the library (one
 * per GPU, held until the process ends; 4 MiB of it for descriptors), so the
 * workspace, sum over the batch of 8*(n*nrhs + 2*64*64*ceil(n/64)) + 8*n + 4
 * bytes (each term rounded up to 256), must fit in 60 MiB: larger batches
 * return HODLR_ERR_SCRATCH.  Results equal hodlr_dgesv_batched's.  Every call
 * on one GPU shares that array: calls that overlap on one GPU (different
 * streams or host threads) race.  Synchronises `stream`. */
int hodlr_dgesv_batched_noalloc(int32_t batch, const int32_t* n, const int32_t* nrhs, double* const* A,
                                const int64_t* lda, double* const* B, const int64_t* ldb, int32_t* info,
                                void* stream);

/* Workspace of system `sys`: N x (c + R) factor [x | A_sub^-1 U] (the
 * reference's U region after the solve), element type 0 = double, 1 = float. */
int hodlr_workspace(hodlr_plan_t plan, int32_t sys, void** ptr, int64_t* ld, int32_t* dtype);

/* Autotuner: times every large GEMM launch of the fused pass and of the
 * correction solve with 64 x 64 and with 128 x 128 workgroup tiles, and every
 * TRSM launch with 64-row and with 128-row steps (`reps` timed rounds after
 * one warm-up round), and keeps the faster choice per launch.  Leaves the
 * workspace in an arbitrary state: factor again after.  `big` (may be NULL)
 * returns how many launches chose the larger tiling. */
int hodlr_autotune(hodlr_plan_t plan, int32_t reps, int32_t* big);
/* Force every tunable launch to the smaller (0) or larger (1) choice; -1
 * restores the built-in defaults. */
int hodlr_tune_force(hodlr_plan_t plan, int32_t cfg);

/* Record the fused pass in a hipGraph and replay it on later calls
 * (needs a non-default stream). */
int hodlr_use_graph(hodlr_plan_t plan, int32_t enable);

/* Per-kernel-category device time: while a category's bit is set in `mask`
 * (bit i = category i, named by hodlr_profile_category(i), i <
 * HODLR_PROFILE_NCAT), every launch of that category is bracketed by HIP
 * events recorded on the plan's stream, without synchronising; read/reset
 * harvest them.  While any bit is set, launches run eagerly, not from a
 * captured graph (event pairs recorded as graph nodes did not bracket their
 * kernels reliably).  Also reports the algorithmic flops and bytes of the
 * timed launches. */
#define HODLR_PROFILE_NCAT 22
int hodlr_profile_enable(hodlr_plan_t plan, int32_t mask);
int hodlr_profile_reset(hodlr_plan_t plan);
int hodlr_profile_read(hodlr_plan_t plan, double* times_ms, int32_t* launches, double* flops, double* bytes);
const char* hodlr_profile_category(int32_t i);

const char* hodlr_error_string(int32_t code);
const char* hodlr_version(void);

/* Device plumbing (no counterpart in the reference, which runs on the CPU):
 * memory, synchronous copies (kind 0 host->device, 1 device->host,
 * 2 device->device), streams and events on the current device. */
int hodlr_device_count(int32_t* n);
int hodlr_device_name(char* buf, int32_t len);
int hodlr_device_malloc(void** ptr, int64_t bytes);
int hodlr_device_free(void* ptr);
int hodlr_memcpy(void* dst, const void* src, int64_t bytes, int32_t kind);
int hodlr_stream_create(void** stream);
int hodlr_stream_destroy(void* stream);
int hodlr_stream_synchronize(void* stream);
int hodlr_device_synchronize(void);
int hodlr_event_create(void** event);
int hodlr_event_destroy(void* event);
int hodlr_event_record(void* event, void* stream);
int hodlr_event_elapsed_ms(void* start, void* stop, double* ms);

#ifdef __cplusplus
}
#endif
#endif

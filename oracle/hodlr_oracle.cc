// ============================================================================
// HODLR CPU oracle -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library, and only as the checker (or as the timed CPU baseline).  The
// GPU product path never calls into it.
//
// It restates, on the CPU and without Legion, the arithmetic of
// fastSolver's HODLR factor-and-solve task bodies, in the reference's order,
// on the reference's column layout:
//   * tree split N/2 | N-N/2 while nrow > threshold .... src/htree/hodlr_matrix.cc:138-169
//   * U region columns [b | U(1) | ... | U(L)] ........... src/htree/hodlr_matrix.cc:44-82,340-352
//   * leaf solve  dgesv_(m, c+R, K, .., X, ..) ........... src/solver/solver_tasks.cc:424-438
//   * reduce      W += V_leaf^T X_leaf[:, range] ......... src/solver/gemm.cc:144-229,315-384
//   * node solve  Schur complement, eta reversed ......... src/solver/solver_tasks.cc:139-305
//   * 2r x 2r direct node form (cross-check only) ....... src/solver/solver_tasks.cc:483-535
//   * broadcast   d -= u * eta ......................... src/solver/gemm.cc:232-284,417-478
//   * schedule    reverse breadth-first visit ........... src/solver/fast_solver.cc:136-255
//   * drand48 rhs, reseeded per region, row-major draws .. src/legion_matrix/init_matrix_tasks.cc:43-77
//   * dense check ||x-x_ref||_F/||x_ref||_F ............. src/solver/direct_solve.cc:216-225
// Additions that the reference does not have (each pinned by an independent
// check in tests/): HODLR matvec y = A x, dense assembly for small N,
// long-double Gaussian elimination, Woodbury closed form for diag*I + U U^T.
//
// Every function is pinned in tests/test_oracle.py (dense LAPACK on the
// assembled matrix, long-double brute force, Woodbury closed form, exact
// integer products, glibc drand48_r, Schur vs 2r x 2r form).
//
// BLAS/LAPACK (dgemm_, dgesv_) are resolved at run time from MKL
// (libmkl_rt.so) or the scipy-bundled OpenBLAS; HODLR_LAPACK overrides.
// ============================================================================
#include <dlfcn.h>
#include <glob.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#include <algorithm>
#include <string>
#include <vector>

namespace {

typedef void (*dgemm_t)(const char*, const char*, const int*, const int*, const int*,
                        const double*, const double*, const int*, const double*, const int*,
                        const double*, double*, const int*);
typedef void (*dgesv_t)(const int*, const int*, double*, const int*, int*, double*,
                        const int*, int*);
typedef void (*set_threads_t)(int);
typedef int (*set_layer_t)(int);

struct Lapack {
  void* h = nullptr;
  dgemm_t dgemm = nullptr;
  dgesv_t dgesv = nullptr;
  set_threads_t set_threads = nullptr;        // MKL_Set_Num_Threads
  set_threads_t set_threads_local = nullptr;  // mkl_set_num_threads_local (returns int; ignored)
  std::string name;
};

Lapack g_lp;

bool try_open(const char* path, const char* prefix, const char* suffix) {
  void* h = dlopen(path, RTLD_NOW | RTLD_GLOBAL);
  if (!h) return false;
  std::string p(prefix), s(suffix);
  void* g = dlsym(h, (p + "dgemm" + s).c_str());
  void* v = dlsym(h, (p + "dgesv" + s).c_str());
  if (!g || !v) { dlclose(h); return false; }
  g_lp.h = h;
  g_lp.dgemm = (dgemm_t)g;
  g_lp.dgesv = (dgesv_t)v;
  g_lp.name = path;
  if (void* lay = dlsym(h, "MKL_Set_Threading_Layer")) ((set_layer_t)lay)(3 /*GNU OpenMP*/);
  g_lp.set_threads = (set_threads_t)dlsym(h, "MKL_Set_Num_Threads");
  g_lp.set_threads_local = (set_threads_t)dlsym(h, "MKL_Set_Num_Threads_Local");
  return true;
}

bool lapack() {
  if (g_lp.h) return true;
  const char* env = getenv("HODLR_LAPACK");
  if (env && try_open(env, "", "_")) return true;
  if (try_open("/opt/conda/lib/libmkl_rt.so", "", "_")) return true;
  if (try_open("libmkl_rt.so", "", "_")) return true;
  if (try_open("libmkl_rt.so.1", "", "_")) return true;
  glob_t gl;
  if (glob("/usr/local/lib/python3*/dist-packages/scipy.libs/libscipy_openblas-*.so", 0, nullptr, &gl) == 0) {
    for (size_t i = 0; i < gl.gl_pathc; ++i)
      if (try_open(gl.gl_pathv[i], "scipy_", "_")) { globfree(&gl); return true; }
    globfree(&gl);
  }
  return false;
}

inline void dgemm(char ta, char tb, int m, int n, int k, double alpha, const double* A, int lda,
                  const double* B, int ldb, double beta, double* C, int ldc) {
  if (m == 0 || n == 0) return;
  g_lp.dgemm(&ta, &tb, &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
}

inline int dgesv(int n, int nrhs, double* A, int lda, double* B, int ldb) {
  std::vector<int> ipiv(std::max(n, 1));
  int info = 0;
  g_lp.dgesv(&n, &nrhs, A, &lda, ipiv.data(), B, &ldb, &info);
  return info;
}

// ---------------------------------------------------------------------------
// Tree: the reference's balanced split, listed breadth first
// (hodlr_matrix.cc:138-169; fast_solver.cc:152-171 builds the same BFS list).
// ---------------------------------------------------------------------------
struct ONode {
  int64_t row_beg, nrow;
  int depth, lc, rc, parent;
};

struct Tree {
  std::vector<ONode> nodes;       // BFS order, nodes[0] = root
  std::vector<int> leaves;        // leaf node ids in row (DFS) order
  std::vector<int64_t> rsum;      // rsum[d] = r_1 + ... + r_d   (rsum[0] = 0)
  int maxdepth = 0;
  int c = 0;
  int rank(int d) const { return int(rsum[d] - rsum[d - 1]); }
  int64_t col_beg(int d) const { return c + rsum[d - 1]; }   // node at depth d >= 1
};

int build_tree(int64_t N, int64_t threshold, int c, int nranks, const int* ranks, Tree& t) {
  t.nodes.clear();
  t.leaves.clear();
  t.c = c;
  t.nodes.push_back({0, N, 0, -1, -1, -1});
  for (size_t i = 0; i < t.nodes.size(); ++i) {
    ONode nd = t.nodes[i];
    if (nd.nrow > threshold) {
      ONode l{nd.row_beg, nd.nrow / 2, nd.depth + 1, -1, -1, int(i)};
      ONode r{nd.row_beg + nd.nrow / 2, nd.nrow - nd.nrow / 2, nd.depth + 1, -1, -1, int(i)};
      t.nodes[i].lc = int(t.nodes.size());
      t.nodes.push_back(l);
      t.nodes[i].rc = int(t.nodes.size());
      t.nodes.push_back(r);
    }
  }
  for (auto& nd : t.nodes) t.maxdepth = std::max(t.maxdepth, nd.depth);
  t.rsum.assign(t.maxdepth + 1, 0);
  for (int d = 1; d <= t.maxdepth; ++d) {
    int r = (nranks == 1) ? ranks[0] : (d - 1 < nranks ? ranks[d - 1] : -1);
    if (r <= 0) return -2;
    t.rsum[d] = t.rsum[d - 1] + r;
  }
  // leaves in row order: DFS
  std::vector<int> stack{0};
  while (!stack.empty()) {
    int i = stack.back();
    stack.pop_back();
    if (t.nodes[i].lc < 0) {
      t.leaves.push_back(i);
      // hodlr_matrix.cc:167 asserts leaf size > rank
      if (t.nodes[i].depth >= 1 && t.nodes[i].nrow <= t.rank(t.nodes[i].depth)) return -3;
    } else {
      stack.push_back(t.nodes[i].rc);
      stack.push_back(t.nodes[i].lc);
    }
  }
  return 0;
}

void leaves_under(const Tree& t, int node, std::vector<int>& out) {
  out.clear();
  std::vector<int> stack{node};
  while (!stack.empty()) {
    int i = stack.back();
    stack.pop_back();
    if (t.nodes[i].lc < 0) out.push_back(i);
    else { stack.push_back(t.nodes[i].rc); stack.push_back(t.nodes[i].lc); }
  }
}

struct Arrays {
  double* X; int64_t ldx;   // N x (c + R): [b | U(1) | ...]
  const double* V; int64_t ldv;   // N x R
  double* K; int64_t ldk; int klayout;  // 0 packed (per leaf m x m, ld m), 1 stacked (row_beg, ld ldk)
};

std::vector<int64_t> leaf_koff(const Tree& t, const Arrays& a) {
  std::vector<int64_t> off(t.nodes.size(), -1);
  int64_t acc = 0;
  for (int li : t.leaves) {
    if (a.klayout == 0) { off[li] = acc; acc += t.nodes[li].nrow * t.nodes[li].nrow; }
    else off[li] = t.nodes[li].row_beg;
  }
  return off;
}

// LUSolveTask::cpu_task (solver_tasks.cc:216-302): Schur complement on r x r.
int node_solve_schur(int r, int nd, const double* V0Tu0, const double* V1Tu1,
                     double* V0Td0, double* V1Td1) {
  // V1Td1 -= V1Tu1 * V0Td0          (solver_tasks.cc:236-240)
  dgemm('n', 'n', r, nd, r, -1.0, V1Tu1, r, V0Td0, r, 1.0, V1Td1, r);
  // S = I - V1Tu1 * V0Tu0           (solver_tasks.cc:254-265)
  std::vector<double> S(size_t(r) * r, 0.0);
  for (int i = 0; i < r; ++i) S[size_t(i) * (r + 1)] = 1.0;
  dgemm('n', 'n', r, r, r, -1.0, V1Tu1, r, V0Tu0, r, 1.0, S.data(), r);
  // S eta0 = V1Td1                  (solver_tasks.cc:278-284)
  int info = dgesv(r, nd, S.data(), r, V1Td1, r);
  if (info != 0) return info;
  // eta1 = V0Td0 - V0Tu0 * eta0     (solver_tasks.cc:298-302)
  dgemm('n', 'n', r, nd, r, -1.0, V0Tu0, r, V1Td1, r, 1.0, V0Td0, r);
  return 0;
}

// serial_leaf_solve interior branch (solver_tasks.cc:483-516): 2r x 2r direct.
int node_solve_direct(int r, int nd, const double* V0Tu0, const double* V1Tu1,
                      double* V0Td0, double* V1Td1) {
  int s = 2 * r;
  std::vector<double> S(size_t(s) * s, 0.0), R(size_t(s) * std::max(nd, 1));
  for (int i = 0; i < s; ++i) S[(r + i) % s + size_t(i) * s] = 1.0;
  for (int j = 0; j < r; ++j)
    for (int i = 0; i < r; ++i) {
      S[i + size_t(j) * s] = V0Tu0[i + size_t(j) * r];
      S[r + i + size_t(r + j) * s] = V1Tu1[i + size_t(j) * r];
    }
  for (int j = 0; j < nd; ++j)
    for (int i = 0; i < r; ++i) {
      R[i + size_t(j) * s] = V0Td0[i + size_t(j) * r];
      R[r + i + size_t(j) * s] = V1Td1[i + size_t(j) * r];
    }
  int info = dgesv(s, nd, S.data(), s, R.data(), s);
  if (info != 0) return info;
  // eta0 = top half, eta1 = bottom half; store in the Schur task's reversed slots
  for (int j = 0; j < nd; ++j)
    for (int i = 0; i < r; ++i) {
      V1Td1[i + size_t(j) * r] = R[i + size_t(j) * s];
      V0Td0[i + size_t(j) * r] = R[r + i + size_t(j) * s];
    }
  return 0;
}

// One interior node visit (fast_solver.cc:191-255).
int visit_node(const Tree& t, int p, const Arrays& a, int node_form, std::vector<int>& scratch) {
  const ONode& P = t.nodes[p];
  int d = P.depth + 1;                // children's depth
  int r = t.rank(d);
  int64_t cb = t.col_beg(d);          // ru = [cb, cb+r), rd = [0, cb)
  int nd = int(cb);
  int64_t vcol = t.rsum[d - 1];       // V columns of depth d
  std::vector<double> V0Tu0(size_t(r) * r, 0.0), V1Tu1(size_t(r) * r, 0.0);
  std::vector<double> V0Td0(size_t(r) * nd, 0.0), V1Td1(size_t(r) * nd, 0.0);
  const int ch[2] = {P.lc, P.rc};
  double* VTu[2] = {V0Tu0.data(), V1Tu1.data()};
  double* VTd[2] = {V0Td0.data(), V1Td1.data()};
  // gemm_reduce x4 (fast_solver.cc:228-235): per legion leaf under the child,
  // W += V_leaf^T X_leaf[:, range]  (gemm.cc:376, beta = 1 into a zeroed W)
  for (int k = 0; k < 2; ++k) {
    leaves_under(t, ch[k], scratch);
    for (int li : scratch) {
      const ONode& L = t.nodes[li];
      const double* Vl = a.V + L.row_beg + vcol * a.ldv;
      const double* Xl = a.X + L.row_beg;
      dgemm('t', 'n', r, r, int(L.nrow), 1.0, Vl, int(a.ldv), Xl + cb * a.ldx, int(a.ldx), 1.0, VTu[k], r);
    }
    for (int li : scratch) {
      const ONode& L = t.nodes[li];
      const double* Vl = a.V + L.row_beg + vcol * a.ldv;
      const double* Xl = a.X + L.row_beg;
      dgemm('t', 'n', r, nd, int(L.nrow), 1.0, Vl, int(a.ldv), Xl, int(a.ldx), 1.0, VTd[k], r);
    }
  }
  int info = node_form == 0 ? node_solve_schur(r, nd, V0Tu0.data(), V1Tu1.data(), V0Td0.data(), V1Td1.data())
                            : node_solve_direct(r, nd, V0Tu0.data(), V1Tu1.data(), V0Td0.data(), V1Td1.data());
  if (info != 0) return info;
  // gemm_broadcast x2 (fast_solver.cc:250-253): d0 -= u0*eta0 (eta0 in V1Td1), d1 -= u1*eta1 (in V0Td0)
  const double* eta[2] = {V1Td1.data(), V0Td0.data()};
  for (int k = 0; k < 2; ++k) {
    leaves_under(t, ch[k], scratch);
    for (int li : scratch) {
      const ONode& L = t.nodes[li];
      double* Xl = a.X + L.row_beg;
      dgemm('n', 'n', int(L.nrow), nd, r, -1.0, Xl + cb * a.ldx, int(a.ldx), eta[k], r, 1.0, Xl, int(a.ldx));
    }
  }
  return 0;
}

int leaf_solve(const Tree& t, int li, const Arrays& a, const std::vector<int64_t>& koff) {
  const ONode& L = t.nodes[li];
  int nrhs = int(L.depth >= 1 ? t.col_beg(L.depth) + t.rank(L.depth) : t.c);
  int64_t ldk = a.klayout == 0 ? L.nrow : a.ldk;
  return dgesv(int(L.nrow), nrhs, a.K + koff[li], int(ldk), a.X + L.row_beg, int(a.ldx));
}

}  // namespace

extern "C" {

int oracle_lapack_name(char* buf, int len) {
  if (!lapack()) return -1;
  snprintf(buf, len, "%s", g_lp.name.c_str());
  return 0;
}

// Version string of the loaded LAPACK (MKL_Get_Version_String, or OpenBLAS's
// get_config); the library path when neither is exported.
int oracle_lapack_version(char* buf, int len) {
  if (!lapack()) return -1;
  typedef void (*mkl_ver_t)(char*, int);
  typedef const char* (*ob_cfg_t)();
  if (void* f = dlsym(g_lp.h, "MKL_Get_Version_String")) {
    ((mkl_ver_t)f)(buf, len);
    return 0;
  }
  for (const char* sym : {"scipy_openblas_get_config64_", "scipy_openblas_get_config", "openblas_get_config"})
    if (void* f = dlsym(g_lp.h, sym)) {
      snprintf(buf, len, "%s", ((ob_cfg_t)f)());
      return 0;
    }
  snprintf(buf, len, "%s", g_lp.name.c_str());
  return 0;
}

// Threads for the BLAS inside one call (MKL only).
void oracle_set_blas_threads(int n) {
  if (lapack() && g_lp.set_threads) g_lp.set_threads(n);
}

// Reference fused factor+solve, in place and destructive like the reference:
// X ([b | U] columns) becomes [x | A_sub^-1 U], K becomes its LU factors.
//   order      0: reverse breadth-first visit, one task at a time (fast_solver.cc:172-181)
//              1: level by level, the independent tasks of a level spread over
//                 OpenMP threads with one BLAS thread each (Legion -ll:cpu N
//                 with sequential BLAS); nthreads <= 0 means all cores.
//   node_form  0: Schur r x r (LUSolveTask); 1: 2r x 2r direct (serial_leaf_solve)
// Returns 0, a LAPACK INFO > 0 of the first failing dgesv, or < 0 on bad input.
int oracle_hodlr_solve(int64_t N, int64_t threshold, int c, int nranks, const int* ranks,
                       double* X, int64_t ldx, const double* V, int64_t ldv,
                       double* K, int64_t ldk, int klayout, int order, int node_form, int nthreads) {
  if (!lapack()) return -1;
  Tree t;
  int rc = build_tree(N, threshold, c, nranks, ranks, t);
  if (rc) return rc;
  Arrays a{X, ldx, V, ldv, K, ldk, klayout};
  std::vector<int64_t> koff = leaf_koff(t, a);
  if (order == 0) {
    std::vector<int> scratch;
    for (int i = int(t.nodes.size()) - 1; i >= 0; --i) {
      int info = (t.nodes[i].lc < 0) ? leaf_solve(t, i, a, koff) : visit_node(t, i, a, node_form, scratch);
      if (info) return info;
    }
    return 0;
  }
  // order 1: leaves (all), then interior nodes by height, tasks of a height in parallel
  if (nthreads <= 0) nthreads = omp_get_num_procs();
  std::vector<int> height(t.nodes.size(), 0);
  int H = 0;
  for (int i = int(t.nodes.size()) - 1; i >= 0; --i)
    if (t.nodes[i].lc >= 0) {
      height[i] = 1 + std::max(height[t.nodes[i].lc], height[t.nodes[i].rc]);
      H = std::max(H, height[i]);
    }
  int status = 0;
  for (int h = 0; h <= H && status == 0; ++h) {
    std::vector<int> tasks;
    for (size_t i = 0; i < t.nodes.size(); ++i) if (height[i] == h) tasks.push_back(int(i));
    int nt = std::min<int>(nthreads, int(tasks.size()));
    if (nt <= 1) {
      if (g_lp.set_threads) g_lp.set_threads(nthreads);
      std::vector<int> scratch;
      for (int i : tasks) {
        int info = h == 0 ? leaf_solve(t, i, a, koff) : visit_node(t, i, a, node_form, scratch);
        if (info) { status = info; break; }
      }
      continue;
    }
#pragma omp parallel num_threads(nt)
    {
      if (g_lp.set_threads_local) g_lp.set_threads_local(1);
      std::vector<int> scratch;
#pragma omp for schedule(dynamic, 1)
      for (int k = 0; k < int(tasks.size()); ++k) {
        int i = tasks[k];
        int info = h == 0 ? leaf_solve(t, i, a, koff) : visit_node(t, i, a, node_form, scratch);
        if (info) {
#pragma omp critical
          if (status == 0) status = info;
        }
      }
      if (g_lp.set_threads_local) g_lp.set_threads_local(0);
    }
  }
  return status;
}

// y = A x for the HODLR matrix given by (V, the U columns of X, K); x, y are N x ncol.
// Not in the reference: the residual of the refinement loop is built on it.
int oracle_hodlr_matvec(int64_t N, int64_t threshold, int c, int nranks, const int* ranks,
                        const double* X, int64_t ldx, const double* V, int64_t ldv,
                        const double* K, int64_t ldk, int klayout,
                        int ncol, const double* x, int64_t ldxx, double* y, int64_t ldy) {
  if (!lapack()) return -1;
  Tree t;
  int rc = build_tree(N, threshold, c, nranks, ranks, t);
  if (rc) return rc;
  Arrays a{const_cast<double*>(X), ldx, V, ldv, const_cast<double*>(K), ldk, klayout};
  std::vector<int64_t> koff = leaf_koff(t, a);
  for (int li : t.leaves) {
    const ONode& L = t.nodes[li];
    int64_t lk = klayout == 0 ? L.nrow : ldk;
    dgemm('n', 'n', int(L.nrow), ncol, int(L.nrow), 1.0, K + koff[li], int(lk), x + L.row_beg, int(ldxx),
          0.0, y + L.row_beg, int(ldy));
  }
  for (const ONode& P : t.nodes) {
    if (P.lc < 0) continue;
    int d = P.depth + 1, r = t.rank(d);
    int64_t ucol = t.col_beg(d), vcol = t.rsum[d - 1];
    const int ch[2] = {P.lc, P.rc};
    std::vector<double> w(size_t(r) * ncol);
    for (int k = 0; k < 2; ++k) {
      const ONode& S = t.nodes[ch[1 - k]];   // sibling supplies V^T x
      const ONode& T = t.nodes[ch[k]];       // this child supplies U
      dgemm('t', 'n', r, ncol, int(S.nrow), 1.0, V + S.row_beg + vcol * ldv, int(ldv), x + S.row_beg, int(ldxx),
            0.0, w.data(), r);
      dgemm('n', 'n', int(T.nrow), ncol, r, 1.0, X + T.row_beg + ucol * ldx, int(ldx), w.data(), r, 1.0,
            y + T.row_beg, int(ldy));
    }
  }
  return 0;
}

// Dense N x N assembly (column-major, ld = N) for small N.
int oracle_hodlr_dense(int64_t N, int64_t threshold, int c, int nranks, const int* ranks,
                       const double* X, int64_t ldx, const double* V, int64_t ldv,
                       const double* K, int64_t ldk, int klayout, double* A) {
  Tree t;
  int rc = build_tree(N, threshold, c, nranks, ranks, t);
  if (rc) return rc;
  Arrays a{const_cast<double*>(X), ldx, V, ldv, const_cast<double*>(K), ldk, klayout};
  std::vector<int64_t> koff = leaf_koff(t, a);
  for (int64_t i = 0; i < N * N; ++i) A[i] = 0.0;
  for (int li : t.leaves) {
    const ONode& L = t.nodes[li];
    int64_t lk = klayout == 0 ? L.nrow : ldk;
    for (int64_t j = 0; j < L.nrow; ++j)
      for (int64_t i = 0; i < L.nrow; ++i)
        A[(L.row_beg + i) + (L.row_beg + j) * N] = K[koff[li] + i + j * lk];
  }
  for (const ONode& P : t.nodes) {
    if (P.lc < 0) continue;
    int d = P.depth + 1, r = t.rank(d);
    int64_t ucol = t.col_beg(d), vcol = t.rsum[d - 1];
    const int ch[2] = {P.lc, P.rc};
    for (int k = 0; k < 2; ++k) {
      const ONode& T = t.nodes[ch[k]];       // rows (U)
      const ONode& S = t.nodes[ch[1 - k]];   // columns (V)
      for (int64_t j = 0; j < S.nrow; ++j)
        for (int64_t i = 0; i < T.nrow; ++i) {
          long double s = 0;
          for (int q = 0; q < r; ++q)
            s += (long double)X[T.row_beg + i + (ucol + q) * ldx] * V[S.row_beg + j + (vcol + q) * ldv];
          A[(T.row_beg + i) + (S.row_beg + j) * N] = double(s);
        }
    }
  }
  return 0;
}

// Dense LAPACK solve of A X = B (A overwritten), as direct_solve.cc:188-195 does.
int oracle_dense_dgesv(int n, int nrhs, double* A, int lda, double* B, int ldb) {
  if (!lapack()) return -1;
  return dgesv(n, nrhs, A, lda, B, ldb);
}

// Long-double Gaussian elimination with partial pivoting (brute force, tiny n).
int oracle_dense_solve_ld(int n, int nrhs, const double* A, int lda, const double* B, int ldb,
                          double* Xo, int ldxo) {
  std::vector<long double> M(size_t(n) * n), R(size_t(n) * nrhs);
  for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) M[i + size_t(j) * n] = A[i + size_t(j) * lda];
  for (int j = 0; j < nrhs; ++j) for (int i = 0; i < n; ++i) R[i + size_t(j) * n] = B[i + size_t(j) * ldb];
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i) if (fabsl(M[i + size_t(k) * n]) > fabsl(M[p + size_t(k) * n])) p = i;
    if (M[p + size_t(k) * n] == 0) return k + 1;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(M[k + size_t(j) * n], M[p + size_t(j) * n]);
      for (int j = 0; j < nrhs; ++j) std::swap(R[k + size_t(j) * n], R[p + size_t(j) * n]);
    }
    for (int i = k + 1; i < n; ++i) {
      long double l = M[i + size_t(k) * n] / M[k + size_t(k) * n];
      if (l == 0) continue;
      for (int j = k; j < n; ++j) M[i + size_t(j) * n] -= l * M[k + size_t(j) * n];
      for (int j = 0; j < nrhs; ++j) R[i + size_t(j) * n] -= l * R[k + size_t(j) * n];
    }
  }
  for (int j = 0; j < nrhs; ++j)
    for (int i = n - 1; i >= 0; --i) {
      long double s = R[i + size_t(j) * n];
      for (int q = i + 1; q < n; ++q) s -= M[i + size_t(q) * n] * R[q + size_t(j) * n];
      R[i + size_t(j) * n] = s / M[i + size_t(i) * n];
    }
  for (int j = 0; j < nrhs; ++j) for (int i = 0; i < n; ++i) Xo[i + size_t(j) * ldxo] = double(R[i + size_t(j) * n]);
  return 0;
}

// Woodbury closed form for the reference's test matrix A = diag*I + U U^T with
// U(i,j) = (i+j) mod r (direct_solve.cc:168-189): x = (b - U (diag*I + U^T U)^-1 U^T b) / diag,
// evaluated in long double.  O(r^3 + N r): U's row i depends on i mod r only.
int oracle_woodbury_circulant(int64_t N, int r, double diag, int nrhs, const double* B, int64_t ldb,
                              double* Xo, int64_t ldxo) {
  // G = U^T U (r x r) = sum over residues p of cnt[p] * u_p u_p^T, u_p(a) =
  // (p + a) mod r, cnt[p] = #{0 <= i < N : i mod r = p}: entries are exact
  // integers (below N r^2 < 2^64, so every partial sum is exact in long double)
  std::vector<long double> G(size_t(r) * r, 0.0L);
  for (int p = 0; p < r && p < N; ++p) {
    const long double cnt = (long double)((N - 1 - p) / r + 1);
    for (int a = 0; a < r; ++a) {
      const long double ua = cnt * (long double)((p + a) % r);
      for (int b = 0; b < r; ++b) G[a + size_t(b) * r] += ua * (long double)((p + b) % r);
    }
  }
  for (int a = 0; a < r; ++a) G[a + size_t(a) * r] += diag;
  for (int j = 0; j < nrhs; ++j) {
    std::vector<long double> z(r, 0.0L);
    for (int64_t i = 0; i < N; ++i)
      for (int a = 0; a < r; ++a) z[a] += (long double)((i + a) % r) * B[i + j * ldb];
    // solve G y = z (G is SPD), Gaussian elimination with partial pivoting
    std::vector<long double> M(G);
    for (int k = 0; k < r; ++k) {
      int p = k;
      for (int i = k + 1; i < r; ++i) if (fabsl(M[i + size_t(k) * r]) > fabsl(M[p + size_t(k) * r])) p = i;
      if (p != k) { for (int q = 0; q < r; ++q) std::swap(M[k + size_t(q) * r], M[p + size_t(q) * r]); std::swap(z[k], z[p]); }
      for (int i = k + 1; i < r; ++i) {
        long double l = M[i + size_t(k) * r] / M[k + size_t(k) * r];
        for (int q = k; q < r; ++q) M[i + size_t(q) * r] -= l * M[k + size_t(q) * r];
        z[i] -= l * z[k];
      }
    }
    for (int i = r - 1; i >= 0; --i) {
      long double s = z[i];
      for (int q = i + 1; q < r; ++q) s -= M[i + size_t(q) * r] * z[q];
      z[i] = s / M[i + size_t(i) * r];
    }
    for (int64_t i = 0; i < N; ++i) {
      long double s = B[i + j * ldb];
      for (int a = 0; a < r; ++a) s -= (long double)((i + a) % r) * z[a];
      Xo[i + j * ldxo] = double(s / diag);
    }
  }
  return 0;
}

// POSIX drand48: X <- (0x5DEECE66D X + 0xB) mod 2^48, value X / 2^48; srand48(seed)
// sets X = ((seed mod 2^32) << 16) + 0x330E.  Fills `nrow` rows x `ncol` columns
// drawing in row-major order, as RandomMatrixTask (init_matrix_tasks.cc:66-76).
void oracle_drand48_rows(int64_t seed, int64_t nrow, int ncol, double* out, int64_t ld) {
  uint64_t x = ((uint64_t(seed) & 0xffffffffULL) << 16) + 0x330EULL;
  const uint64_t a = 0x5DEECE66DULL, mask = (1ULL << 48) - 1;
  for (int64_t i = 0; i < nrow; ++i)
    for (int j = 0; j < ncol; ++j) {
      x = (a * x + 0xBULL) & mask;
      out[i + j * ld] = ldexp(double(x), -48);
    }
}

// glibc's drand48_r, for pinning the restatement above.
void oracle_glibc_drand48_rows(int64_t seed, int64_t nrow, int ncol, double* out, int64_t ld) {
  struct drand48_data buf;
  srand48_r(long(seed), &buf);
  for (int64_t i = 0; i < nrow; ++i)
    for (int j = 0; j < ncol; ++j) drand48_r(&buf, &out[i + j * ld]);
}

// Relative error of direct_solve.cc:216-225: sqrt(sum (x - xref)^2 / sum xref^2).
double oracle_rel_l2(int64_t n, int ncol, const double* x, int64_t ldx, const double* xref, int64_t ldr) {
  long double diff = 0, den = 0;
  for (int j = 0; j < ncol; ++j)
    for (int64_t i = 0; i < n; ++i) {
      long double d = (long double)x[i + j * ldx] - xref[i + j * ldr];
      diff += d * d;
      den += (long double)xref[i + j * ldr] * xref[i + j * ldr];
    }
  return double(sqrtl(diff / den));
}

}  // extern "C"

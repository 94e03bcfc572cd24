// C++ facade of fastsolver_amd: the reference's driver API (HodlrMatrix,
// FastSolver, Range) without the Legion Context / HighLevelRuntime arguments,
// over the C-ABI of libfastsolver_amd.so and libfastsolver_amd_gen.so.
//
//   reference                                         here
//   HodlrMatrix(col,row,gl,sl,r,t,leaf,name)          same  (include/htree/hodlr_matrix.h:19-21)
//   create_tree(ctx, rt)                              create_tree()          (hodlr_matrix.h:24-25)
//   init_rhs(seed, range, ctx, rt)                    init_rhs(seed, range)  (hodlr_matrix.h:26-27)
//   init_circulant_matrix(diag, range, ctx, rt, skipU) init_circulant_matrix(diag, range, skipU)
//   save_rhs / save_solution                          same (text format of SaveRegionTask)
//   launch_level, get_num_rhs, get_num_leaf, get_file_soln   same
//   display_launch_time()                             same: host time of init_rhs + init_circulant_matrix
//   Node* uroot, vroot (hodlr_matrix.h:49-50)         same names; Node (include/htree/node.h:12-43) keeps
//                                                     nrow/ncol/row_beg/col_beg/lchild/rchild, and the
//                                                     node's device block (ptr, ld) replaces its LMatrix*
//   FastSolver::bfs_solve(hmat, range, ctx, rt)       bfs_solve(hmat, range) (fast_solver.h:19-20)
//   register_solver_tasks()                           register_solver_tasks(): checks a GPU is present
//   compute_L2_error(hmat, seed, rows, nregions,      same minus ctx/rt (include/solver/direct_solve.h:20-24)
//                    rhs_cols, rank, diag, ctx, rt)
//
// The matrix lives in device memory owned by the HodlrMatrix; the solve is
// not destructive (the reference overwrites K and U, hodlr_matrix.cc / solver_tasks.cc:437).
#ifndef FASTSOLVER_AMD_HPP
#define FASTSOLVER_AMD_HPP

#include <stdint.h>
#include <stdlib.h>

#include <chrono>
#include <cmath>
#include <memory>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fastsolver_amd.h"
#include "fastsolver_amd_gen.h"

namespace fastsolver_amd {

inline void check(int rc, const char* what) {
  if (rc != HODLR_OK) throw std::runtime_error(std::string(what) + ": " + hodlr_error_string(rc));
}

// Placement hint of the reference (include/core/range.h); one GPU ignores it.
class Range {
 public:
  Range() : mbegin(0), msize(0) {}
  explicit Range(int size) : mbegin(0), msize(size) {}
  Range(int begin, int size) : mbegin(begin), msize(size) {}
  int begin() const { return mbegin; }
  int size() const { return msize; }
  Range lchild() const { return Range(mbegin, msize / 2); }
  Range rchild() const { return Range(mbegin + msize / 2, msize / 2); }
  int mbegin, msize;
};

// Tree node of the reference (include/htree/node.h:12-43).  The U tree has
// the reference's column layout (root: the c right-hand-side columns; a child
// of a depth-(d-1) node: `rank` columns at col_beg = c + (d-1)*rank of X); the
// V tree's blocks sit at col_beg = (d-1)*rank of V.  Where the reference keeps
// an LMatrix (a Legion region) per node, `ptr` points at the node's nrow x ncol
// block in device memory (leading dimension `ld`).  legion leaf = real leaf.
struct Node {
  int nrow = 0, ncol = 0, row_beg = 0, col_beg = 0;
  Node* lchild = nullptr;
  Node* rchild = nullptr;
  double* ptr = nullptr;
  int64_t ld = 0;
  bool is_real_leaf() const { return lchild == nullptr && rchild == nullptr; }
  bool is_legion_leaf() const { return is_real_leaf(); }
};

// Device buffer owned by the facade.
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(int64_t bytes) { check(hodlr_device_malloc(&p_, bytes), "hodlr_device_malloc"); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_) { o.p_ = nullptr; }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    release();
    p_ = o.p_;
    o.p_ = nullptr;
    return *this;
  }
  ~DeviceBuffer() { release(); }
  void* get() const { return p_; }
  void upload(const void* h, int64_t bytes) { check(hodlr_memcpy(p_, h, bytes, 0), "upload"); }
  void download(void* h, int64_t bytes) const { check(hodlr_memcpy(h, p_, bytes, 1), "download"); }

 private:
  void release() {
    if (p_) hodlr_device_free(p_);
    p_ = nullptr;
  }
  void* p_ = nullptr;
};

class HodlrMatrix {
 public:
  HodlrMatrix() = default;
  HodlrMatrix(int col, int row, int gl, int sl, int r, int t, int leaf, const std::string& name)
      : rhs_cols(col), rhs_rows(row), gloLevel(gl), subLevel(sl), rank(r), threshold(t), leafSize(leaf),
        file_rhs(name + "_rhs.txt"), file_soln(name + "_soln.txt") {}

  // Builds the tree (hodlr_matrix.cc:44-82) and allocates the device arrays.
  void create_tree() {
    int nleaves = 0, maxdepth = 0;
    int64_t xcols = 0, vcols = 0, kpacked = 0;
    int32_t r = rank;
    check(hodlr_gen_tree(rhs_rows, threshold, rhs_cols, 1, &r, &nleaves, &maxdepth, &xcols, &vcols, &kpacked),
          "create_tree");
    nLegionLeaf = nleaves;
    depth = maxdepth;
    ncolX = xcols;
    ncolV = vcols;
    nK = kpacked;
    X = DeviceBuffer(int64_t(rhs_rows) * xcols * 8);
    V = DeviceBuffer(int64_t(rhs_rows) * std::max<int64_t>(vcols, 1) * 8);
    K = DeviceBuffer(kpacked * 8);
    x = DeviceBuffer(int64_t(rhs_rows) * rhs_cols * 8);
    nodes.clear();
    uroot = new_node(rhs_rows, rhs_cols, 0, 0, (double*)X.get());
    vroot = new_node(rhs_rows, 0, 0, 0, (double*)V.get());
    split(uroot, vroot, 1);
  }

  // The drand48 right-hand side, reseeded with `seed` at the start of every
  // leaf region and drawn row by row (hodlr_matrix.cc:175-209,
  // init_matrix_tasks.cc:43-77): generated on the host and written into the
  // first rhs_cols columns of X.  Independent of init_circulant_matrix, so the
  // two calls may come in either order.
  void init_rhs(long seed, const Range& = Range()) {
    const auto t0 = std::chrono::steady_clock::now();
    rhs_seed = seed;
    std::vector<double> hb(size_t(rhs_rows) * rhs_cols);
    check(hodlr_gen_rhs_drand48(rhs_rows, threshold, rhs_cols, rank, seed, 0, hb.data(), rhs_rows), "init_rhs");
    X.upload(hb.data(), int64_t(hb.size()) * 8);   // ld = rhs_rows: the first rhs_cols columns are contiguous
    timeInit += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // A = diag*I + U U^T with U(i,j) = (i+j) mod rank (hodlr_matrix.cc:94-109):
  // U into X's columns rhs_cols.. (skipped with skipU, as the reference's
  // init_Umat), V, and the leaf blocks K.  The right-hand-side columns are
  // left as init_rhs wrote them.
  void init_circulant_matrix(double diag, const Range& = Range(), bool skipU = false) {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<double> hX(size_t(rhs_rows) * ncolX);
    std::vector<double> hV(size_t(rhs_rows) * std::max<int64_t>(ncolV, 1));
    std::vector<double> hK(static_cast<size_t>(nK));
    check(hodlr_gen_circulant(rhs_rows, threshold, rhs_cols, rank, diag, rhs_seed, 0, hX.data(), rhs_rows,
                              hV.data(), rhs_rows, hK.data(), 0, HODLR_K_PACKED),
          "init_circulant_matrix");
    const size_t off = size_t(rhs_rows) * rhs_cols;
    if (!skipU)
      check(hodlr_memcpy((double*)X.get() + off, hX.data() + off, int64_t(hX.size() - off) * 8, 0),
            "init_circulant_matrix");
    V.upload(hV.data(), int64_t(hV.size()) * 8);
    K.upload(hK.data(), int64_t(hK.size()) * 8);
    timeInit += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // hodlr_matrix.h:43-46: the time spent initialising the matrix and the
  // right-hand side (host generation and upload here; task launches there).
  void display_launch_time() const {
    std::cout << "Time cost for launching init-tasks :" << timeInit << " s" << std::endl;
  }

  // Tree roots (hodlr_matrix.h:49-50), valid after create_tree.
  Node* uroot = nullptr;
  Node* vroot = nullptr;

  // Text dumps in SaveRegionTask's format (save_region_task.cc:37-67): per leaf
  // region, optionally the seed, then one row per line of tab-terminated values.
  void save_rhs() const { save(file_rhs, true); }
  void save_solution() const { save(file_soln, false); }

  int launch_level() const { return gloLevel - subLevel; }
  int get_num_rhs() const { return rhs_cols; }
  int get_num_leaf() const { return nLegionLeaf; }
  std::string get_file_soln() const { return file_soln; }
  int rows() const { return rhs_rows; }
  int levels() const { return depth; }

  hodlr_system_t system() const {
    hodlr_system_t s{};
    s.n = rhs_rows;
    s.threshold = threshold;
    s.nrhs = rhs_cols;
    s.nranks = 1;
    s.ranks = &rank;
    s.X = (const double*)X.get();
    s.ldx = rhs_rows;
    s.V = (const double*)V.get();
    s.ldv = rhs_rows;
    s.K = (const double*)K.get();
    s.klayout = HODLR_K_PACKED;
    s.x = (double*)x.get();
    s.ldxs = rhs_rows;
    return s;
  }

  // solution x (rows x rhs columns, column-major)
  std::vector<double> solution() const {
    std::vector<double> h(size_t(rhs_rows) * rhs_cols);
    x.download(h.data(), int64_t(h.size()) * 8);
    return h;
  }
  std::vector<double> rhs() const {
    std::vector<double> h(size_t(rhs_rows) * rhs_cols);
    X.download(h.data(), int64_t(h.size()) * 8);
    return h;
  }

 private:
  void save(const std::string& file, bool seed) const {
    std::vector<double> h = seed ? rhs() : solution();
    std::vector<int64_t> rb(nLegionLeaf), nr(nLegionLeaf);
    std::vector<int> dp(nLegionLeaf);
    int32_t r = rank;
    check(hodlr_gen_leaves(rhs_rows, threshold, rhs_cols, 1, &r, rb.data(), nr.data(), dp.data()), "save");
    std::ofstream out(file.c_str());
    for (int l = 0; l < nLegionLeaf; ++l) {
      if (seed) out << rhs_seed << std::endl;
      for (int64_t i = rb[l]; i < rb[l] + nr[l]; ++i) {
        for (int j = 0; j < rhs_cols; ++j) out << std::setprecision(20) << h[i + size_t(j) * rhs_rows] << '\t';
        out << std::endl;
      }
    }
  }

  Node* new_node(int nrow, int ncol, int row_beg, int col_beg, double* base) {
    nodes.emplace_back(new Node());
    Node* p = nodes.back().get();
    p->nrow = nrow;
    p->ncol = ncol;
    p->row_beg = row_beg;
    p->col_beg = col_beg;
    p->ld = rhs_rows;
    p->ptr = base + row_beg + int64_t(col_beg) * rhs_rows;
    return p;
  }
  // split rule of create_balanced_tree (hodlr_matrix.cc:138-169): N/2 | N - N/2
  // while nrow > threshold; children of a depth-(d-1) node sit at depth d
  void split(Node* u, Node* v, int d) {
    if (u->nrow <= threshold) return;
    const int n0 = u->nrow / 2, n1 = u->nrow - n0;
    const int ucol = u->col_beg + u->ncol, vcol = (d - 1) * rank;
    u->lchild = new_node(n0, rank, u->row_beg, ucol, (double*)X.get());
    u->rchild = new_node(n1, rank, u->row_beg + n0, ucol, (double*)X.get());
    v->lchild = new_node(n0, rank, v->row_beg, vcol, (double*)V.get());
    v->rchild = new_node(n1, rank, v->row_beg + n0, vcol, (double*)V.get());
    split(u->lchild, v->lchild, d + 1);
    split(u->rchild, v->rchild, d + 1);
  }

  std::vector<std::unique_ptr<Node>> nodes;
  double timeInit = 0;
  int rhs_cols = 0, rhs_rows = 0, gloLevel = 0, subLevel = 0;
  int32_t rank = 0;
  int threshold = 0, leafSize = 1, nLegionLeaf = 0, depth = 0;
  int64_t ncolX = 0, ncolV = 0, nK = 0;
  long rhs_seed = 0;
  std::string file_rhs, file_soln;
  DeviceBuffer X, V, K, x;
};

inline void register_solver_tasks() {
  int32_t n = 0;
  check(hodlr_device_count(&n), "register_solver_tasks");
  if (n < 1) throw std::runtime_error("register_solver_tasks: no GPU visible");
}

class FastSolver {
 public:
  explicit FastSolver(int precision = HODLR_PREC_FP64) : precision_(precision) {}
  FastSolver(const FastSolver&) = delete;
  FastSolver& operator=(const FastSolver&) = delete;
  ~FastSolver() { release(); }

  // Factor + solve (fast_solver.cc:56-72); FP32 plans add FP64 refinement to `tol`.
  void bfs_solve(HodlrMatrix& m, const Range& = Range(), double tol = 1e-12, int maxiter = 20) {
    release();
    check(hodlr_stream_create(&stream_), "hodlr_stream_create");
    hodlr_system_t s = m.system();
    check(hodlr_plan_create(&plan_, &s, 1, precision_, stream_), "hodlr_plan_create");
    hodlr_event_create(&t0_);
    hodlr_event_create(&t1_);
    hodlr_event_record(t0_, stream_);
    if (precision_ == HODLR_PREC_FP64) {
      check(hodlr_factor_solve(plan_), "hodlr_factor_solve");
    } else {
      check(hodlr_solve(plan_, tol, maxiter, &report_), "hodlr_solve");
    }
    hodlr_event_record(t1_, stream_);
    check(hodlr_stream_synchronize(stream_), "bfs_solve");
    hodlr_event_elapsed_ms(t0_, t1_, &time_ms_);
    int32_t nn = 0;
    int rc = hodlr_status(plan_, nullptr, nullptr, &nn);
    check(rc, "bfs_solve status");
  }

  // max over rhs columns of ||b - A x|| / ||b||
  double residual() {
    double mx = 0;
    check(hodlr_residual(plan_, nullptr, &mx), "hodlr_residual");
    return mx;
  }

  void display_launch_time() const {
    std::cout << "Time for the GPU factor+solve : " << time_ms_ / 1e3 << " s" << std::endl;
  }
  double elapsed_ms() const { return time_ms_; }
  const hodlr_solve_report_t& report() const { return report_; }

 private:
  void release() {
    if (plan_) hodlr_plan_destroy(plan_);
    if (t0_) hodlr_event_destroy(t0_);
    if (t1_) hodlr_event_destroy(t1_);
    if (stream_) hodlr_stream_destroy(stream_);
    plan_ = nullptr;
    t0_ = t1_ = stream_ = nullptr;
  }
  int precision_;
  hodlr_plan_t plan_ = nullptr;
  void *stream_ = nullptr, *t0_ = nullptr, *t1_ = nullptr;
  double time_ms_ = 0;
  hodlr_solve_report_t report_{};
};

// The reference's error report for its circulant test matrix
// (src/solver/direct_solve.cc:128-245): saves the solver's solution,
// regenerates the right-hand side (drand48 reseeded with `seed` at the start of
// each of `nregions` row blocks, drawn row by row), solves A = diag*I + U U^T,
// U(i,j) = (i+j) mod rank, exactly -- by the Woodbury identity in long double,
// x = (b - U (diag*I + U^T U)^-1 U^T b) / diag, where the reference runs a dense
// N x N dgesv -- writes it to soln_ref.txt, prints and returns
// sqrt(sum (x - x_ref)^2 / sum x_ref^2).
inline double compute_L2_error(const HodlrMatrix& m, long seed, int rhs_rows, int nregions, int rhs_cols,
                               int rank, double diag) {
  m.save_solution();
  const int n = rhs_rows, c = rhs_cols, r = rank;
  std::vector<double> b(size_t(n) * c);
  const int block = n / nregions;
  for (int g = 0; g < nregions; ++g) {
    drand48_data buf;
    srand48_r(seed, &buf);
    for (int i = 0; i < block; ++i)
      for (int j = 0; j < c; ++j) drand48_r(&buf, &b[size_t(g * block + i) + size_t(j) * n]);
  }
  typedef long double LD;
  auto U = [r](int i, int k) { return LD((i + k) % r); };
  // G = diag*I + U^T U (r x r), W = U^T b (r x c)
  std::vector<LD> G(size_t(r) * r, 0), W(size_t(r) * c, 0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < r; ++k) {
      const LD u = U(i, k);
      for (int l = 0; l < r; ++l) G[k + size_t(l) * r] += u * U(i, l);
      for (int j = 0; j < c; ++j) W[k + size_t(j) * r] += u * LD(b[i + size_t(j) * n]);
    }
  for (int k = 0; k < r; ++k) G[k + size_t(k) * r] += diag;
  // Gaussian elimination with partial pivoting: W <- G^-1 W
  for (int k = 0; k < r; ++k) {
    int p = k;
    for (int i = k + 1; i < r; ++i)
      if (std::fabs(double(G[i + size_t(k) * r])) > std::fabs(double(G[p + size_t(k) * r]))) p = i;
    if (p != k) {
      for (int l = 0; l < r; ++l) std::swap(G[k + size_t(l) * r], G[p + size_t(l) * r]);
      for (int j = 0; j < c; ++j) std::swap(W[k + size_t(j) * r], W[p + size_t(j) * r]);
    }
    for (int i = k + 1; i < r; ++i) {
      const LD f = G[i + size_t(k) * r] / G[k + size_t(k) * r];
      for (int l = k; l < r; ++l) G[i + size_t(l) * r] -= f * G[k + size_t(l) * r];
      for (int j = 0; j < c; ++j) W[i + size_t(j) * r] -= f * W[k + size_t(j) * r];
    }
  }
  for (int j = 0; j < c; ++j)
    for (int k = r - 1; k >= 0; --k) {
      LD x = W[k + size_t(j) * r];
      for (int l = k + 1; l < r; ++l) x -= G[k + size_t(l) * r] * W[l + size_t(j) * r];
      W[k + size_t(j) * r] = x / G[k + size_t(k) * r];
    }
  std::vector<double> xref(size_t(n) * c);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < c; ++j) {
      LD acc = b[i + size_t(j) * n];
      for (int k = 0; k < r; ++k) acc -= U(i, k) * W[k + size_t(j) * r];
      xref[i + size_t(j) * n] = double(acc / diag);
    }
  {
    std::ofstream ofs("soln_ref.txt");
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < c; ++j) ofs << std::setprecision(20) << xref[i + size_t(j) * n] << '\t';
      ofs << std::endl;
    }
  }
  std::vector<double> soln(size_t(n) * c, 0.0);
  {
    std::ifstream ifs(m.get_file_soln().c_str());
    if (!ifs) throw std::runtime_error("compute_L2_error: cannot open " + m.get_file_soln());
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < c; ++j)
        if (!(ifs >> soln[i + size_t(j) * n])) throw std::runtime_error("compute_L2_error: short solution file");
  }
  double diff = 0, denom = 0;
  for (size_t e = 0; e < soln.size(); ++e) {
    diff += (soln[e] - xref[e]) * (soln[e] - xref[e]);
    denom += xref[e] * xref[e];
  }
  const double err = std::sqrt(diff / denom);
  std::cout << "Err: " << err << std::endl;
  return err;
}

}  // namespace fastsolver_amd

#endif

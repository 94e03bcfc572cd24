// Host-side input generation for HODLR systems (not on the solve path).
//
// Two matrix families:
//  * the reference's own test matrix A = diag*I + U U^T with U(i,j) = (i+j) mod r
//    and a drand48 right-hand side reseeded per leaf region, drawn row by row
//    (src/legion_matrix/init_matrix_tasks.cc:43-77,176-219;
//     src/htree/hodlr_matrix.cc:191-266,476-567);
//  * seeded random families: counter-based Philox4x32-10 uniforms (Salmon et
//    al., SC'11) and Box-Muller normals.  Every (system, kind, node, element)
//    has its own counter, so any subset can be generated independently and in
//    parallel, and the same host arrays feed the CPU and the GPU runs.
//
// The decaying-spectrum family orthonormalises Gaussian blocks with LAPACK
// dgeqrf/dorgqr, resolved at run time (MKL libmkl_rt, else scipy's OpenBLAS).
#include <dlfcn.h>
#include <glob.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../tree.hpp"
#include "../../../include/fastsolver_amd_gen.h"

namespace {

// ---------------------------------------------------------------- Philox4x32-10
inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  uint64_t p = uint64_t(a) * uint64_t(b);
  hi = uint32_t(p >> 32);
  lo = uint32_t(p);
}

inline void philox(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(M0, c0, hi0, lo0);
    mulhilo(M1, c2, hi1, lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += W0; k1 += W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

enum Kind : uint32_t { KIND_RHS = 1, KIND_U = 2, KIND_V = 3, KIND_K = 4 };

struct Stream {
  uint32_t key[2];
  uint32_t node, tag;
  Stream(uint64_t seed, uint32_t system, Kind kind, uint32_t node_id) {
    key[0] = uint32_t(seed);
    key[1] = uint32_t(seed >> 32);
    node = node_id;
    tag = (system << 4) | uint32_t(kind);
  }
  // the two 53-bit uniforms in [0,1) of counter block `blk`
  inline void uniforms(uint64_t blk, double u[2]) const {
    uint32_t in[4] = {uint32_t(blk), uint32_t(blk >> 32), node, tag}, o[4];
    philox(in, key, o);
    u[0] = double((uint64_t(o[0]) << 21) | (o[1] >> 11)) * 0x1.0p-53;
    u[1] = double((uint64_t(o[2]) << 21) | (o[3] >> 11)) * 0x1.0p-53;
  }
  inline double uniform(uint64_t e) const {
    double u[2];
    uniforms(e >> 1, u);
    return u[e & 1];
  }
  inline double normal(uint64_t e) const {
    double u[2];
    uniforms(e >> 1, u);
    double rad = sqrt(-2.0 * log(1.0 - u[0]));
    double ang = 6.283185307179586476925286766559 * u[1];
    return (e & 1) ? rad * sin(ang) : rad * cos(ang);
  }
};

// ---------------------------------------------------------------- LAPACK (QR)
typedef void (*dgeqrf_t)(const int*, const int*, double*, const int*, double*, double*, const int*, int*);
typedef void (*dorgqr_t)(const int*, const int*, const int*, double*, const int*, const double*, double*,
                         const int*, int*);
typedef int (*mkl_local_t)(int);
struct Qr {
  void* h = nullptr;
  dgeqrf_t geqrf = nullptr;
  dorgqr_t orgqr = nullptr;
  mkl_local_t local_threads = nullptr;
  std::string name;
} g_qr;

bool qr_open(const char* path, const char* prefix) {
  void* h = dlopen(path, RTLD_NOW | RTLD_GLOBAL);
  if (!h) return false;
  std::string p(prefix);
  g_qr.geqrf = (dgeqrf_t)dlsym(h, (p + "dgeqrf_").c_str());
  g_qr.orgqr = (dorgqr_t)dlsym(h, (p + "dorgqr_").c_str());
  if (!g_qr.geqrf || !g_qr.orgqr) { dlclose(h); return false; }
  g_qr.h = h;
  g_qr.name = path;
  if (void* lay = dlsym(h, "MKL_Set_Threading_Layer")) ((int (*)(int))lay)(3 /*GNU OpenMP*/);
  g_qr.local_threads = (mkl_local_t)dlsym(h, "MKL_Set_Num_Threads_Local");
  return true;
}

bool qr_lib() {
  if (g_qr.h) return true;
  const char* env = getenv("HODLR_LAPACK");
  if (env && qr_open(env, "")) return true;
  if (qr_open("/opt/conda/lib/libmkl_rt.so", "") || qr_open("libmkl_rt.so", "") || qr_open("libmkl_rt.so.1", ""))
    return true;
  glob_t gl;
  bool ok = false;
  if (glob("/usr/local/lib/python3*/dist-packages/scipy.libs/libscipy_openblas-*.so", 0, nullptr, &gl) == 0) {
    for (size_t i = 0; i < gl.gl_pathc && !ok; ++i) ok = qr_open(gl.gl_pathv[i], "scipy_");
    globfree(&gl);
  }
  return ok;
}

// Q (n x r, column-major, ld n) from the thin Householder QR of Q's contents.
int thin_q(int n, int r, double* Q) {
  int info = 0, lwork = -1;
  std::vector<double> tau(r);
  double wq = 0;
  g_qr.geqrf(&n, &r, Q, &n, tau.data(), &wq, &lwork, &info);
  lwork = std::max(int(wq), r * 64);
  std::vector<double> work(lwork);
  g_qr.geqrf(&n, &r, Q, &n, tau.data(), work.data(), &lwork, &info);
  if (info) return info;
  g_qr.orgqr(&n, &r, &r, Q, &n, tau.data(), work.data(), &lwork, &info);
  return info;
}

std::vector<int64_t> leaf_koffsets(const hodlr::Tree& t, int klayout) {
  std::vector<int64_t> off(t.leaves.size());
  int64_t acc = 0;
  for (size_t j = 0; j < t.leaves.size(); ++j) {
    const hodlr::TreeNode& x = t.nodes[t.leaves[j]];
    off[j] = klayout == HODLR_K_STACKED ? x.row_beg : acc;
    acc += x.nrow * x.nrow;
  }
  return off;
}

}  // namespace

extern "C" {

void hodlr_gen_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox(ctr, key, out);
}

int hodlr_gen_stream_values(uint64_t seed, uint32_t system, uint32_t kind, uint32_t node, int normal,
                            int64_t first, int64_t count, double* out) {
  Stream s(seed, system, Kind(kind), node);
  for (int64_t e = 0; e < count; ++e) out[e] = normal ? s.normal(uint64_t(first + e)) : s.uniform(uint64_t(first + e));
  return 0;
}

int hodlr_gen_lapack_name(char* buf, int len) {
  if (!qr_lib()) return -1;
  snprintf(buf, len, "%s", g_qr.name.c_str());
  return 0;
}

}  // extern "C"

namespace {
// Right-hand side of the reference: drand48 reseeded with `seed` at the start
// of every region, values drawn row by row into the column-major block
// (init_matrix_tasks.cc:67-76); the regions are the leaves
// (hodlr_matrix.cc:191-209) or blocks of rhs_region_rows rows.
void rhs_drand48(const hodlr::Tree& t, int64_t seed, int64_t rhs_region_rows, double* X, int64_t ldx) {
  const int64_t n = t.n;
  const int nrhs = t.nrhs;
  std::vector<std::pair<int64_t, int64_t>> regions;
  if (rhs_region_rows > 0) {
    for (int64_t b = 0; b < n; b += rhs_region_rows) regions.push_back({b, std::min(rhs_region_rows, n - b)});
  } else {
    for (int id : t.leaves) regions.push_back({t.nodes[id].row_beg, t.nodes[id].nrow});
  }
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < int64_t(regions.size()); ++g) {
    uint64_t x = ((uint64_t(seed) & 0xffffffffULL) << 16) + 0x330EULL;
    for (int64_t i = 0; i < regions[g].second; ++i)
      for (int j = 0; j < nrhs; ++j) {
        x = (0x5DEECE66DULL * x + 0xBULL) & ((1ULL << 48) - 1);
        X[regions[g].first + i + j * ldx] = ldexp(double(x), -48);
      }
  }
}
}  // namespace

extern "C" {

int hodlr_gen_rhs_drand48(int64_t n, int64_t threshold, int nrhs, int rank, int64_t seed,
                          int64_t rhs_region_rows, double* X, int64_t ldx) {
  hodlr::Tree t;
  int rc = hodlr::build_tree(n, threshold, nrhs, 1, &rank, t);
  if (rc) return rc;
  rhs_drand48(t, seed, rhs_region_rows, X, ldx);
  return 0;
}

int hodlr_gen_circulant(int64_t n, int64_t threshold, int nrhs, int rank, double diag, int64_t seed,
                        int64_t rhs_region_rows, double* X, int64_t ldx, double* V, int64_t ldv, double* K,
                        int64_t ldk, int klayout) {
  hodlr::Tree t;
  int rc = hodlr::build_tree(n, threshold, nrhs, 1, &rank, t);
  if (rc) return rc;
  const int64_t r = rank;
  rhs_drand48(t, seed, rhs_region_rows, X, ldx);
  // U and V: value (i_global + j) mod r in every level block
  // (init_matrix_tasks.cc:208-218; hodlr_matrix.cc:476-497)
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < t.vcols(); ++j)
    for (int64_t i = 0; i < n; ++i) {
      double v = double((i + j) % r);
      X[i + (nrhs + j) * ldx] = v;
      V[i + j * ldv] = v;
    }
  // leaf blocks K = diag*I + U_leaf U_leaf^T, U_leaf(i,q) = (row_beg+i+q) mod r
  // (hodlr_matrix.cc:528-567); all entries are exact integers plus diag
  std::vector<int64_t> koff = leaf_koffsets(t, klayout);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t lj = 0; lj < int64_t(t.leaves.size()); ++lj) {
    const hodlr::TreeNode& x = t.nodes[t.leaves[lj]];
    int64_t lk = klayout == HODLR_K_STACKED ? ldk : x.nrow;
    double* Kl = K + koff[lj];
    for (int64_t b = 0; b < x.nrow; ++b)
      for (int64_t a = 0; a < x.nrow; ++a) {
        int64_t s = 0;
        for (int64_t q = 0; q < r; ++q) s += ((x.row_beg + a + q) % r) * ((x.row_beg + b + q) % r);
        Kl[a + b * lk] = double(s) + (a == b ? diag : 0.0);
      }
  }
  return 0;
}

int hodlr_gen_random(int64_t n, int64_t threshold, int nrhs, int nranks, const int* ranks,
                     const hodlr_family* fam, double* X, int64_t ldx, double* V, int64_t ldv, double* K,
                     int64_t ldk, int klayout) {
  hodlr::Tree t;
  int rc = hodlr::build_tree(n, threshold, nrhs, nranks, ranks, t);
  if (rc) return rc;
  if (fam->uv_mode == HODLR_UV_DECAY && !qr_lib()) return -201;
  const uint64_t seed = fam->seed;
  const uint32_t sys = fam->system_id;
  // right-hand side uniform[0,1)
  {
    Stream s(seed, sys, KIND_RHS, 0);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < nrhs; ++j)
      for (int64_t i = 0; i < n; ++i) X[i + j * ldx] = s.uniform(uint64_t(i + j * n));
  }
  // dense leaf blocks D = randn * scale + shift * I
  std::vector<int64_t> koff = leaf_koffsets(t, klayout);
  for (size_t lj = 0; lj < t.leaves.size(); ++lj) {
    const hodlr::TreeNode& x = t.nodes[t.leaves[lj]];
    int64_t lk = klayout == HODLR_K_STACKED ? ldk : x.nrow;
    double scale = fam->d_scale > 0 ? fam->d_scale : 1.0 / sqrt(double(x.nrow));
    Stream s(seed, sys, KIND_K, uint32_t(t.leaves[lj]));
    double* Kl = K + koff[lj];
    const int64_t m = x.nrow;
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < m; ++b)
      for (int64_t a = 0; a < m; ++a)
        Kl[a + b * lk] = s.normal(uint64_t(a + b * m)) * scale + (a == b ? fam->d_shift : 0.0);
  }
  // low-rank factors of every non-root node
  std::vector<int> order;
  for (size_t i = 1; i < t.nodes.size(); ++i) order.push_back(int(i));
  int status = 0;
  if (fam->uv_mode == HODLR_UV_UNIFORM) {
    for (int id : order) {
      const hodlr::TreeNode& x = t.nodes[id];
      const int r = t.rank(x.depth);
      const int64_t ucol = t.col_beg(x.depth), vcol = t.rsum[x.depth - 1];
      const double sc = 1.0 / sqrt(double(x.nrow));
      Stream su(seed, sys, KIND_U, uint32_t(id)), sv(seed, sys, KIND_V, uint32_t(id));
#pragma omp parallel for schedule(static)
      for (int64_t q = 0; q < r; ++q)
        for (int64_t a = 0; a < x.nrow; ++a) {
          uint64_t e = uint64_t(a + q * x.nrow);
          X[x.row_beg + a + (ucol + q) * ldx] = (2.0 * su.uniform(e) - 1.0) * sc;
          V[x.row_beg + a + (vcol + q) * ldv] = (2.0 * sv.uniform(e) - 1.0) * sc;
        }
    }
    return 0;
  }
  // decaying spectrum: U = Q_U diag(10^(-k/decay)), V = Q_V.  Small blocks are
  // orthonormalised in parallel (one BLAS thread each), big ones one at a time.
  const int nthr = omp_get_max_threads();
  auto one = [&](int id, int which, bool parallel) -> int {
    const hodlr::TreeNode& x = t.nodes[id];
    const int r = t.rank(x.depth);
    const int nn = int(x.nrow);
    Stream s(seed, sys, which == 0 ? KIND_U : KIND_V, uint32_t(id));
    std::vector<double> Q(size_t(nn) * r);
    const int64_t ne = int64_t(Q.size());
    if (parallel) {
#pragma omp parallel for schedule(static)
      for (int64_t e = 0; e < ne; ++e) Q[e] = s.normal(uint64_t(e));
    } else {
      for (int64_t e = 0; e < ne; ++e) Q[e] = s.normal(uint64_t(e));
    }
    int info = thin_q(nn, r, Q.data());
    if (info) return info;
    if (which == 0) {
      const int64_t ucol = t.col_beg(x.depth);
#pragma omp parallel for schedule(static) if (parallel)
      for (int q = 0; q < r; ++q) {
        double sig = pow(10.0, -double(q) / fam->decay);
        for (int a = 0; a < nn; ++a) X[x.row_beg + a + (ucol + q) * ldx] = Q[a + size_t(q) * nn] * sig;
      }
    } else {
      const int64_t vcol = t.rsum[x.depth - 1];
#pragma omp parallel for schedule(static) if (parallel)
      for (int q = 0; q < r; ++q)
        for (int a = 0; a < nn; ++a) V[x.row_beg + a + (vcol + q) * ldv] = Q[a + size_t(q) * nn];
    }
    return 0;
  };
  std::vector<int> small, big;
  for (int id : order) (t.nodes[id].nrow * t.rank(t.nodes[id].depth) < (int64_t(1) << 22) ? small : big).push_back(id);
  for (int id : big)
    for (int w = 0; w < 2; ++w) { int info = one(id, w, true); if (info && !status) status = info; }
#pragma omp parallel num_threads(nthr)
  {
    if (g_qr.local_threads) g_qr.local_threads(1);
#pragma omp for schedule(dynamic, 1)
    for (int64_t k = 0; k < int64_t(small.size()) * 2; ++k) {
      int info = one(small[k / 2], int(k % 2), false);
      if (info) {
#pragma omp critical
        if (!status) status = info;
      }
    }
    if (g_qr.local_threads) g_qr.local_threads(0);
  }
  return status;
}

// Tree query for hosts that need the layout (leaf row ranges, depths, K offsets).
int hodlr_gen_tree(int64_t n, int64_t threshold, int nrhs, int nranks, const int* ranks, int* nleaves,
                   int* maxdepth, int64_t* xcols, int64_t* vcols, int64_t* kpacked) {
  hodlr::Tree t;
  int rc = hodlr::build_tree(n, threshold, nrhs, nranks, ranks, t);
  if (rc) return rc;
  *nleaves = int(t.leaves.size());
  *maxdepth = t.maxdepth;
  *xcols = t.xcols();
  *vcols = t.vcols();
  int64_t acc = 0;
  for (int id : t.leaves) acc += t.nodes[id].nrow * t.nodes[id].nrow;
  *kpacked = acc;
  return 0;
}

int hodlr_gen_leaves(int64_t n, int64_t threshold, int nrhs, int nranks, const int* ranks, int64_t* row_beg,
                     int64_t* nrow, int* depth) {
  hodlr::Tree t;
  int rc = hodlr::build_tree(n, threshold, nrhs, nranks, ranks, t);
  if (rc) return rc;
  for (size_t j = 0; j < t.leaves.size(); ++j) {
    row_beg[j] = t.nodes[t.leaves[j]].row_beg;
    nrow[j] = t.nodes[t.leaves[j]].nrow;
    depth[j] = t.nodes[t.leaves[j]].depth;
  }
  return 0;
}

}  // extern "C"

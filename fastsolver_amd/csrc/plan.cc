// Plan, breadth-first level schedule and C-ABI of the HODLR solver.
//
// The reference launches one Legion task per leaf, per reduce, per node solve
// and per broadcast, walking its breadth-first node list backwards
// (src/solver/fast_solver.cc:136-255).  Here every phase of every node of the
// same height, across all systems of the plan, is one batched kernel launch
// over a descriptor array built once at plan creation:
//
//   leaf      K copy -> LU (recursive, 512-column steps of 64-wide panels: panel /
//             inverse of its diagonal block / U rows as MFMA products / trailing
//             update) -> perm -> TRSM of [b | U] (solver_tasks.cc:424-438)
//   height h  reduce   W_i = V_i^T X_i[:, 0:c+R_l+r]  (split-K + fixed-order sum)
//             node     VTd1 -= VTu1 VTd0;  S = I - VTu1 VTu0;  LU(S);  eta0 = S^-1 VTd1;
//                      eta1 = VTd0 - VTu0 eta0          (solver_tasks.cc:216-302)
//             bcast    X_i[:, 0:c+R_l] -= X_i[:, ucols] eta_i  (gemm.cc:417-478)
//
// Heights replace the reference's reverse BFS order: all children of a node of
// height h have height < h, so one launch per phase and height respects every
// dependence Legion derived from region privileges.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fastsolver_amd.h"
#include "desc.hpp"
#include "kernels.hpp"
#include "tree.hpp"

namespace hodlr {
namespace {

constexpr int kMaxBlock = 1024;   // largest leaf / rank the LU panel kernel holds in registers

enum Cat {
  CAT_LEAF_COPY = 0, CAT_LU_PANEL, CAT_LU_ROW, CAT_LU_UPDATE, CAT_LU_MISC, CAT_LEAF_TRSM,
  CAT_REDUCE, CAT_REDUCE_SUM, CAT_NODE_GEMM, CAT_NODE_LU, CAT_NODE_TRSM, CAT_BCAST,
  CAT_COPY, CAT_MATVEC, CAT_NORM, CAT_OTHER, CAT_SOLVE_TRSM, CAT_SOLVE_REDUCE, CAT_SOLVE_NODE, CAT_SOLVE_BCAST,
  CAT_MATVEC_VTX, CAT_MATVEC_ROWS,
  CAT_N
};
// matvec: the sums and gathers of the HODLR product; matvec_vtx: its V^T x
// products; matvec_rows: its blockdiag(K) x + U Z products
const char* kCatNames[CAT_N] = {"leaf_copy", "lu_panel", "lu_rowblock", "lu_update", "lu_misc", "leaf_trsm",
                                "reduce_gemm", "reduce_sum", "node_gemm", "node_lu", "node_trsm", "bcast_gemm",
                                "copy", "matvec", "norm", "other", "solve_trsm", "solve_reduce", "solve_node",
                                "solve_bcast", "matvec_vtx", "matvec_rows"};
static_assert(CAT_N == HODLR_PROFILE_NCAT, "category count");

enum Kind { K_GEMM, K_SUM, K_COPY, K_LU_INIT, K_LU_PANEL, K_LU_PERM, K_LU_DINV, K_LU_SMALL, K_TRSM, K_NORM, K_NORM_FIN, K_TRANS,
            K_CHECK, K_STREAM, K_FRAG };

struct Launch {
  Kind kind;
  Cat cat;
  DT t[4];
  bool trans = false;
  int role = 0;
  int tile = 0;          // GEMM workgroup tile: 0 = 64 x 64, 1 = 128 x 128, 2 = 64 x 32, 3 = row panel,
                         // 4 = bf16 reduce (kReduceRows x kReduceCols)
  int gid = -1;          // tunable launches: index into the plan's choices
  int alt[2] = {0, 1};   // tunable launches: the tile of choice 0 and of choice 1
  size_t off_t[5] = {0, 0, 0, 0, 0};
  int64_t tiles_t[5] = {0, 0, 0, 0, 0};
  size_t off = 0;        // byte offset of the first descriptor in the blob
  int count = 0;
  int64_t tiles = 0;
  int k0 = 0, rows_max = 0, mmax = 0;
  int mode = 0;          // K_CHECK: 0 after the first residual, 1 after a sweep; K_LU_DINV: blocks
                         // from k0 / kLuNB to invert (0: one)
  int nmax = 0;          // K_STREAM: the widest product of the batch (columns)
  bool gated = false;    // a refinement-sweep launch: runs only while the plan's RefineCtl::active is set
  double flops = 0, bytes = 0;   // algorithmic work of the launch (profiling)
};

struct Proc {
  std::vector<Launch> l;
  void append(const Proc& o) { l.insert(l.end(), o.l.begin(), o.l.end()); }
};

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline size_t esize(DT t) { return t == F64 ? 8 : t == F32 ? 4 : 2; }   // storage bytes


// Device buffer with typed offsets.
struct Buf {
  char* p = nullptr;
  size_t bytes = 0;
};

struct SysState {
  hodlr_system_t s;
  Tree tree;
  std::vector<int32_t> ranks;
  int64_t ldw = 0;                 // leading dimension of the workspaces (= N)
  size_t xws = 0, yws = 0, rbuf = 0, lu = 0, lpiv = 0, lperm = 0, linfo = 0, zbuf = 0, ldinv = 0;
  size_t vws = 0;                  // FP32 plans: V rounded to FP32 once per factor pass (N x R, ld N)
  std::vector<size_t> vb;          // bf16 plans, per non-root node (bytes into ws): its V block (nrow x r)
                                   //   rounded to bf16, as MFMA A fragments of V^T (FragDesc layout)
  size_t mvx = 0;                  // hodlr_matvec: the caller's x staged here (N x c FP64, ld N)
  std::vector<size_t> lu_off;      // per leaf (elements)
  std::vector<size_t> dinv_off;    // per leaf (elements): inverted 64 x 64 diagonal blocks
  size_t lut = 0;                  // bf16 plans: transposed bf16 copy of the leaf LU factors
  std::vector<size_t> lut_off;     // per leaf (elements), ld = m rounded up to 32
  std::vector<size_t> z_off;       // per leaf (doubles)
  std::vector<size_t> vtd, vtu, wmv;   // per node (doubles): reduce outputs
  std::vector<int> leaf_lo, leaf_hi;   // per node: its leaves are leaf_id leaf_lo .. leaf_hi-1
  size_t pmv = 0;                  // matvec: per leaf V_j^T x_j, pld x c each (doubles)
  int64_t pld = 0;
  std::vector<size_t> sbuf, eta, spiv, sperm, sdinv;  // per interior node
  std::vector<size_t> etaf;        // per interior node, FP32 plans: eta0 | eta1 in the workspace type
  std::vector<int> sinfo;          // per interior node: index into node info array
  std::vector<size_t> kofs;        // per leaf (elements of K)
  size_t bnorm = 0, rnorm = 0;     // doubles, c each
};

}  // namespace
}  // namespace hodlr

using namespace hodlr;

struct hodlr_plan_s {
  int precision = HODLR_PREC_FP64;
  DT wt = F64;                     // workspace element type
  DT mt = F64;                     // math type of the leaf updates, TRSM updates, reduce and broadcast
  hipStream_t stream = nullptr;
  std::vector<SysState> sys;
  int height = 0, nleaves = 0, nnodes = 0;
  // device memory
  Buf ws;        // all workspaces (one allocation)
  Buf descs;     // all descriptors
  Buf scratch;   // split-K partials and norm partials
  size_t scratch_doubles = 0;
  int32_t* leaf_info = nullptr;    // into ws
  int32_t* node_info = nullptr;
  // procedures
  Proc leaf, extract, residual, update;
  Proc matvec;       // rbuf = A mvx (hodlr_matvec)
  Proc solve_out;    // rbuf = yws (FP64), the result of a correction solve
  std::vector<Proc> reduce, node, bcast;          // index = height
  std::vector<Proc> reduce_s, node_s, bcast_s;    // solve-only (c columns)
  Proc solve_only_leaf;
  Proc factor;                                    // everything of the fused pass
  // device-driven refinement (hodlr_solve): `first` = residual of the factor
  // pass's solution + convergence check; `sweep` = correction solve, x += y,
  // residual, check, every launch gated by RefineCtl::active
  Proc first, sweep;
  size_t ctl = 0;                  // RefineCtl, byte offset in ws
  RefineCtl ctl_host{};            // host copy (written before a solve, read after)
  int sweep_hint = 0;              // sweeps the last hodlr_solve took
  // profiling: event pairs around the launches of the selected categories,
  // recorded asynchronously and harvested by hodlr_profile_read
  uint32_t prof_mask = 0;
  double prof_ms[CAT_N] = {0};
  int32_t prof_n[CAT_N] = {0};
  double prof_flops[CAT_N] = {0}, prof_bytes[CAT_N] = {0};
  struct Pending { int cat; hipEvent_t a, b; double flops, bytes; };
  std::vector<Pending> pending;
  // hip graphs: the fused pass (hodlr_factor_solve); a whole solve (fused pass,
  // first residual and `sweeps` gated sweeps); one sweep.  Not used while a
  // profile mask is set.
  struct Graph {
    hipGraphExec_t exec = nullptr;
    int sweeps = 0;
  };
  bool use_graph = false;
  Graph g_factor, g_solve, g_sweep;
  bool graphs() const { return use_graph && prof_mask == 0; }
  std::vector<hipEvent_t> pool;
  // GEMM tilings: per tunable launch the chosen tile (0: 64 x 64, 1: 128 x 128)
  std::vector<int> tile_choice;
  bool tuning = false;
  int tune_cfg = 0;
  std::vector<double> tune_ms[2];
  struct TunePending { int gid; hipEvent_t a, b; };
  std::vector<TunePending> tune_pending;
  // info
  hodlr_plan_info_t info{};
  bool factored = false;
};

namespace {

// ------------------------------------------------------------ descriptor blob
struct Blob {
  std::vector<char> data;
  template <class D> size_t add(const std::vector<D>& v) {
    size_t off = (data.size() + 255) & ~size_t(255);
    data.resize(off + v.size() * sizeof(D));
    memcpy(data.data() + off, v.data(), v.size() * sizeof(D));
    return off;
  }
};

struct Builder {
  hodlr_plan_s* P;
  Blob blob;
  size_t scratch_need = 0;   // doubles
  size_t ngid = 0;           // GEMM launches with both tilings
  bool tunable = false;      // the plan's own launches (one-off batches are not tuned)
  bool bad_gate = false;     // a refinement sweep holds a launch kind without a gate

  char* wsp(size_t off) { return P->ws.p + off; }

  // ---- GEMM batches
  struct G {
    const void* A; int64_t lda; const void* B; int64_t ldb; const void* Cin; int64_t ldcin; void* Cout;
    int64_t ldcout; int m, n, k; double alpha, beta;
  };
  // split: desired minimum number of workgroups; 0 = never split
  // min_wg: split the K range of beta = 0 problems into chunks of equal row
  // count until the batch has about min_wg workgroups (0 = never split).
  void gemm(Proc& pr, Cat cat, DT ta, DT tb, DT tc, DT mt, bool trans, int role, const std::vector<G>& gs,
            int64_t min_wg = 0) {
    std::vector<GemmDesc> d;
    std::vector<SumDesc> sums;
    // tile size: 64 x 64 (measured the faster one for nearly every launch of
    // the cfg-family workloads); launches of roles 1-3 also carry the 128 x 128
    // tiling, which the plan's autotuner may switch to.  Launches whose
    // problems are all at most 32 columns wide (the residual and correction
    // sweeps on c <= 32 right-hand sides) use 64 x 32 tiles.
    int tile = 0;
    bool narrow = true;
    for (const G& g : gs)
      if (g.m > 0 && g.n > 0) narrow = narrow && g.n <= kGemmNarrowBN;
    if (narrow) tile = 2;
    double work = 0;   // 64-tile rows of K work (independent of the tiling)
    for (const G& g : gs)
      if (g.m > 0 && g.n > 0) work += double(cdiv(g.m, kGemmBM) * cdiv(g.n, kGemmBN)) * std::max(g.k, 1);
    const int64_t kMinChunk = 1024;
    const int64_t chunk = min_wg > 0 ? std::max<int64_t>(kMinChunk, int64_t(work / double(min_wg))) : 0;
    size_t part = 0;  // doubles of split-K partials
    double flops = 0, bytes = 0;
    for (const G& g : gs) {
      if (g.m <= 0 || g.n <= 0) continue;
      GemmDesc x{};
      x.A = g.A; x.B = g.B; x.Cin = g.Cin; x.Cout = g.Cout;
      x.lda = g.lda; x.ldb = g.ldb; x.ldcin = g.ldcin; x.ldcout = g.ldcout;
      x.m = g.m; x.n = g.n; x.k = g.k;
      x.alpha = g.alpha; x.beta = g.beta;
      int nsplit = 1;
      if (chunk > 0 && g.beta == 0.0 && tc == F64 && g.k >= 2 * chunk) nsplit = int(cdiv(g.k, chunk));
      int kchunk = int(cdiv(cdiv(std::max(g.k, 1), nsplit), kGemmBK) * kGemmBK);
      nsplit = int(cdiv(std::max(g.k, 1), kchunk));
      x.kchunk = kchunk;
      x.nsplit = nsplit;
      if (nsplit > 1) {
        // partials go to scratch (doubles: every split-K caller outputs FP64)
        x.Cout = (void*)(uintptr_t)(part * 8);   // patched to the scratch base below
        x.ldcout = g.m;
        x.split_stride = int64_t(g.m) * g.n;
        SumDesc sd{};
        sd.P = (const void*)(uintptr_t)(part * 8);
        sd.stride = x.split_stride;
        sd.out = g.Cout;
        sd.ldo = g.ldcout;
        sd.m = g.m; sd.n = g.n; sd.nsplit = nsplit; sd.ldp = g.m;
        sums.push_back(sd);
        part += size_t(nsplit) * g.m * g.n;
      }
      d.push_back(x);
      flops += 2.0 * g.m * g.n * g.k;
      bytes += double(g.m) * g.k * esize(ta) + double(g.k) * g.n * esize(tb) +
               double(g.m) * g.n * esize(tc) * (g.beta != 0.0 ? 2 : 1);
    }
    if (d.empty()) return;
    if (part) {
      scratch_need = std::max(scratch_need, part);
      for (auto& x : d)
        if (x.nsplit > 1) x.Cout = (void*)(P->scratch.p + (uintptr_t)x.Cout);
      for (auto& s : sums) s.P = (const void*)(P->scratch.p + (uintptr_t)s.P);
    }
    Launch L;
    L.kind = K_GEMM; L.cat = cat;
    L.t[0] = ta; L.t[1] = tb; L.t[2] = tc; L.t[3] = mt;
    L.trans = trans; L.role = role; L.tile = tile;
    L.count = int(d.size());
    L.flops = flops; L.bytes = bytes;
    // broadcasts (role 2, every k <= 128, no split) also get the row-panel
    // kernel, the default for them; the engine tiling stays the alternative
    int kmax = 0;
    for (auto& x : d) kmax = std::max(kmax, x.k);
    bool panel = role == 2 && kmax <= kPanelKMax && ta == F32 && tb == F32 && tc == F32 && mt == BF16;
    for (auto& x : d) panel = panel && x.ldb % 8 == 0 && x.ldb >= x.k;
    for (int cfg = 0; cfg < 4; ++cfg) {
      if (cfg == 3) {
        if (!panel) continue;
        int64_t tiles = 0;
        for (auto& x : d) {
          x.tiles_m = int(cdiv(x.m, 64));
          x.tiles_n = 1;
          x.tile_begin = tiles;
          tiles += x.tiles_m;
        }
        L.off_t[3] = blob.add(d);
        L.tiles_t[3] = tiles;
        continue;
      }
      if (narrow ? cfg != 2 : (cfg == 2 || (cfg == 1 && (role == 0 || role == 4)))) continue;
      const int TM = cfg == 1 ? 128 : kGemmBM, TN = cfg == 1 ? 128 : (cfg == 2 ? kGemmNarrowBN : kGemmBN);
      int64_t tiles = 0;
      for (auto& x : d) {
        x.tiles_m = int(cdiv(x.m, TM));
        x.tiles_n = int(cdiv(x.n, TN));
        x.tile_begin = tiles;
        tiles += int64_t(x.tiles_m) * x.tiles_n * x.nsplit;
      }
      L.off_t[cfg] = blob.add(d);
      L.tiles_t[cfg] = tiles;
    }
    if (panel) {
      L.tile = 3;
      L.alt[0] = 3;
      L.alt[1] = narrow ? 2 : 0;
      if (tunable) L.gid = int(ngid++);
    } else if (role >= 1 && role <= 3 && tunable && !narrow) {
      L.gid = int(ngid++);
    }
    L.off = L.off_t[L.tile];
    L.tiles = L.tiles_t[L.tile];
    pr.l.push_back(L);
    sum(pr, cat == CAT_REDUCE ? CAT_REDUCE_SUM : cat, sums);
  }

  // bf16 plans' broadcasts with eta as bf16 MFMA fragments (G::ldb = their
  // column-group count): the row-panel kernel only, 64-row panels
  void bcast_frag(Proc& pr, Cat cat, const std::vector<G>& gs) {
    std::vector<GemmDesc> d;
    double flops = 0, bytes = 0;
    int64_t tiles = 0;
    for (const G& g : gs) {
      if (g.m <= 0 || g.n <= 0) continue;
      GemmDesc x{};
      x.A = g.A; x.B = g.B; x.Cin = g.Cin; x.Cout = g.Cout;
      x.lda = g.lda; x.ldb = g.ldb; x.ldcin = g.ldcin; x.ldcout = g.ldcout;
      x.m = g.m; x.n = g.n; x.k = g.k; x.kchunk = g.k; x.nsplit = 1;
      x.alpha = g.alpha; x.beta = g.beta;
      x.tiles_m = int(cdiv(g.m, 64));
      x.tiles_n = 1;
      x.tile_begin = tiles;
      tiles += x.tiles_m;
      d.push_back(x);
      flops += 2.0 * g.m * g.n * g.k;
      bytes += double(g.m) * g.k * 4 + double(g.k) * g.n * 2 + double(g.m) * g.n * 4 * (g.beta != 0.0 ? 2 : 1);
    }
    if (d.empty()) return;
    Launch L;
    L.kind = K_GEMM; L.cat = cat;
    L.t[0] = L.t[1] = L.t[2] = F32; L.t[3] = BF16;
    L.role = 2; L.tile = 3; L.alt[0] = L.alt[1] = 3; L.mode = 1;
    L.count = int(d.size()); L.flops = flops; L.bytes = bytes;
    L.off = L.off_t[3] = blob.add(d);
    L.tiles = L.tiles_t[3] = tiles;
    pr.l.push_back(L);
  }

  void sum(Proc& pr, Cat cat, std::vector<SumDesc> v) {
    std::vector<SumDesc> d;
    int64_t st = 0;
    for (auto& s : v) {
      if (s.m <= 0 || s.n <= 0 || s.nsplit <= 0) continue;
      s.tile_begin = st;
      st += cdiv(int64_t(s.m) * s.n, kSumTile);
      d.push_back(s);
    }
    if (d.empty()) return;
    Launch S;
    S.kind = K_SUM; S.cat = cat;
    S.off = blob.add(d); S.count = int(d.size()); S.tiles = st;
    pr.l.push_back(S);
  }

  void copy(Proc& pr, Cat cat, DT ts, DT td, std::vector<CopyDesc> v) {
    int64_t t = 0;
    std::vector<CopyDesc> d;
    double bytes = 0;
    for (auto& c : v) {
      if (c.rows <= 0 || c.cols <= 0) continue;
      c.tile_begin = t;
      t += cdiv(int64_t(c.rows) * c.cols, kCopyTile);
      bytes += double(c.rows) * c.cols * (esize(ts) + esize(td) * (c.beta != 0.0 ? 2 : 1));
      d.push_back(c);
    }
    if (d.empty()) return;
    Launch L;
    L.kind = K_COPY; L.cat = cat; L.t[0] = ts; L.t[1] = td;
    L.off = blob.add(d); L.count = int(d.size()); L.tiles = t; L.bytes = bytes;
    pr.l.push_back(L);
  }

  // Blocked LU with partial pivoting of a batch of square blocks.
  void lu(Proc& pr, Cat panel_cat, Cat row_cat, Cat upd_cat, Cat misc_cat, DT t, const std::vector<LuDesc>& v,
          DT upd_mt = F64) {
    if (upd_mt != BF16) upd_mt = t;
    if (v.empty()) return;
    int mmax = 0;
    for (auto& x : v) mmax = std::max(mmax, x.m);
    size_t off = blob.add(v);
    if (mmax <= kSmallLU) {
      // every block in registers: one launch for the whole factorization
      // (info included), then the inverted diagonal blocks and perm; a
      // 256-thread workgroup per block of 65-128 rows, one wave per block of
      // at most 64 (batches of thousands of them -- many small systems --
      // keep many blocks in flight per CU)
      Launch S;
      S.kind = K_LU_SMALL; S.cat = panel_cat; S.t[0] = t; S.off = off; S.count = int(v.size()); S.mmax = mmax;
      for (auto& x : v) S.flops += 2.0 / 3.0 * double(x.m) * x.m * x.m;
      pr.l.push_back(S);
      Launch Dv;
      Dv.kind = K_LU_DINV; Dv.cat = misc_cat; Dv.t[0] = t; Dv.off = off; Dv.count = int(v.size()); Dv.k0 = 0;
      Dv.mode = int(cdiv(mmax, kLuNB));
      pr.l.push_back(Dv);
      Launch Pm;
      Pm.kind = K_LU_PERM; Pm.cat = misc_cat; Pm.off = off; Pm.count = int(v.size());
      pr.l.push_back(Pm);
      return;
    }
    Launch I;
    I.kind = K_LU_INIT; I.cat = misc_cat; I.off = off; I.count = int(v.size());
    pr.l.push_back(I);
    const size_t es = esize(t);
    // factor the 64 columns at kp and invert their diagonal block (final once
    // the panel is done; the row solves and the TRSMs use it)
    auto panel = [&](int kp) {
      Launch Pn;
      Pn.kind = K_LU_PANEL; Pn.cat = panel_cat; Pn.t[0] = t; Pn.off = off; Pn.count = int(v.size());
      Pn.k0 = kp; Pn.rows_max = mmax - kp;
      double pf = 0;
      for (auto& x : v)
        if (x.m > kp) { int nb = std::min(kLuNB, x.m - kp); pf += double(x.m - kp) * nb * nb; }
      Pn.flops = pf;
      pr.l.push_back(Pn);
      Launch Dv;
      Dv.kind = K_LU_DINV; Dv.cat = misc_cat; Dv.t[0] = t; Dv.off = off; Dv.count = int(v.size()); Dv.k0 = kp;
      pr.l.push_back(Dv);
    };
    // U12 = L_kp^-1 A(kp : kp+64, c0 : c0+n) in place, n = min(cmax, m - c0):
    // one 64-row tile per column block, so every workgroup reads its columns
    // before it overwrites them
    auto row_solve = [&](int kp, int c0, int cmax) {
      std::vector<G> rs;
      for (auto& x : v) {
        const int n = std::min(cmax, x.m - c0);
        if (n <= 0) continue;
        char* A = (char*)x.A;
        G g;
        g.A = (char*)x.dinv + es * size_t(kp / kLuNB) * 2 * kLuNB * kLuNB;
        g.lda = kLuNB;
        g.B = A + es * (size_t(kp) + size_t(c0) * x.lda);
        g.ldb = x.lda;
        g.Cout = (void*)g.B;
        g.Cin = g.Cout;
        g.ldcin = g.ldcout = x.lda;
        g.m = kLuNB; g.n = n; g.k = kLuNB;
        g.alpha = 1.0; g.beta = 0.0;
        rs.push_back(g);
      }
      gemm(pr, row_cat, t, t, t, t, false, 0, rs);
    };
    // A(r0 : r0+mr, c0 : c0+n) -= A(r0 : r0+mr, kc : kc+kk) A(kc : kc+kk, c0 : c0+n)
    auto update = [&](int r0, int rmax, int c0, int cmax, int kc, int kk) {
      std::vector<G> gs;
      for (auto& x : v) {
        const int mr = std::min(rmax, x.m - r0), n = std::min(cmax, x.m - c0);
        if (mr <= 0 || n <= 0) continue;
        char* A = (char*)x.A;
        G g;
        g.A = A + es * (size_t(r0) + size_t(kc) * x.lda);
        g.lda = x.lda;
        g.B = A + es * (size_t(kc) + size_t(c0) * x.lda);
        g.ldb = x.lda;
        g.Cout = A + es * (size_t(r0) + size_t(c0) * x.lda);
        g.Cin = g.Cout;
        g.ldcin = g.ldcout = x.lda;
        g.m = mr; g.n = n; g.k = kk;
        g.alpha = -1.0; g.beta = 1.0;
        gs.push_back(g);
      }
      gemm(pr, upd_cat, t, t, t, upd_mt, false, 3, gs);
    };
    // Recursive LU in steps of kLuStep columns: a step factors its columns as
    // two halves (down to 64-column panels), and the trailing matrix is read
    // and written once per step with K = kLuStep.  Row interchanges commute
    // with the deferred row updates, since each panel swaps whole rows of
    // every column outside it.
    constexpr int kBig = 1 << 30;
    // rows kp .. kp+h of columns c0 .. c0+cmax: L(kp : kp+h, kp : kp+h)^-1 times them
    std::function<void(int, int, int, int)> trsm_rows = [&](int kp, int h, int c0, int cmax) {
      if (h == kLuNB) { row_solve(kp, c0, cmax); return; }
      trsm_rows(kp, h / 2, c0, cmax);
      update(kp + h / 2, h / 2, c0, cmax, kp, h / 2);
      trsm_rows(kp + h / 2, h / 2, c0, cmax);
    };
    // factor columns kp .. kp+w (every row below kp), earlier updates applied
    std::function<void(int, int)> factor = [&](int kp, int w) {
      if (w == kLuNB) { panel(kp); return; }
      const int h = w / 2;
      factor(kp, h);
      if (mmax <= kp + h) return;
      trsm_rows(kp, h, kp + h, h);                  // the right half's rows of U
      update(kp + h, kBig, kp + h, h, kp, h);       // the right half below them
      factor(kp + h, h);
    };
    for (int k0 = 0; k0 < mmax; k0 += kLuStep) {
      factor(k0, kLuStep);
      const int k2 = k0 + kLuStep;
      if (mmax <= k2) break;
      trsm_rows(k0, kLuStep, k2, kBig);
      update(k2, kBig, k2, kBig, k0, kLuStep);     // trailing matrix, K = kLuStep
    }
    Launch Pm;
    Pm.kind = K_LU_PERM; Pm.cat = misc_cat; Pm.off = off; Pm.count = int(v.size());
    pr.l.push_back(Pm);
  }

  // Both TRSM variants (64-row steps / 128-row steps) get descriptor sets;
  // the plan's launches are tunable like the GEMM tilings.
  void trsm(Proc& pr, Cat cat, DT tlu, DT tin, DT tout, std::vector<TrsmDesc> v, DT ut = F64) {
    if (ut != BF16) ut = tlu;
    std::vector<TrsmDesc> d;
    double flops = 0, bytes = 0;
    int mmax = 0;
    for (auto& x : v) {
      if (x.m <= 0 || x.ncols <= 0) continue;
      flops += 2.0 * double(x.m) * x.m * x.ncols;
      bytes += double(x.m) * x.m * esize(tlu) + double(x.m) * x.ncols * (esize(tin) + esize(tout));
      mmax = std::max(mmax, x.m);
      d.push_back(x);
    }
    if (d.empty()) return;
    Launch L;
    L.kind = K_TRSM; L.cat = cat; L.t[0] = tlu; L.t[1] = tin; L.t[2] = tout; L.t[3] = ut;
    L.count = int(d.size()); L.flops = flops; L.bytes = bytes; L.mmax = mmax;
    // choices 0 / 1: 64-row / 128-row steps; bf16 plans: LDS-resident solved
    // rows with 64 / 32 columns per workgroup (variants 2 / 3)
    for (int cfg = 0; cfg < 2; ++cfg) {
      const int var = ut == BF16 ? 2 + cfg : cfg;
      int64_t t = 0;
      for (auto& x : d) {
        x.tile_begin = t;
        t += cdiv(x.ncols, trsm_panel_cols(tlu, var));
      }
      L.off_t[cfg] = blob.add(d);
      L.tiles_t[cfg] = t;
    }
    L.tile = 0;   // the faster choice on every measured workload
    if (tunable) L.gid = int(ngid++);
    L.off = L.off_t[L.tile];
    L.tiles = L.tiles_t[L.tile];
    pr.l.push_back(L);
  }

  // bf16 MFMA A fragments of transposed fp64 blocks (bf16 plans' V copies)
  void frag(Proc& pr, Cat cat, std::vector<FragDesc> v) {
    int64_t t = 0;
    double bytes = 0;
    bool vec = true;   // 32-byte loads: aligned sources, whole 8-row slots
    for (auto& x : v) {
      x.tile_begin = t;
      const int64_t slots = cdiv(x.rows, 32) * x.ngroups * 64;
      t += x.ngroups * cdiv(cdiv(x.rows, 32), 4 * kFragPerWave);   // column group x row-block run
      bytes += double(x.rows) * x.cols * 8 + double(slots) * 16;
      vec = vec && (uintptr_t)x.src % 32 == 0 && x.lds % 4 == 0 && x.rows % 8 == 0;
    }
    if (v.empty()) return;
    Launch L;
    L.kind = K_FRAG; L.cat = cat; L.off = blob.add(v); L.count = int(v.size()); L.tiles = t; L.bytes = bytes;
    L.mode = vec;
    pr.l.push_back(L);
  }

  // transposed bf16 copies of LU factors (bf16 plans)
  void transpose(Proc& pr, Cat cat, std::vector<TransDesc> v) {
    int64_t t = 0;
    double bytes = 0;
    for (auto& x : v) {
      x.tile_begin = t;
      t += cdiv(x.m, 64) * x.tiles_k;
      bytes += double(x.m) * x.m * 4 + double(x.m) * x.ldd * 2;
    }
    if (v.empty()) return;
    Launch L;
    L.kind = K_TRANS; L.cat = cat; L.off = blob.add(v); L.count = int(v.size()); L.tiles = t; L.bytes = bytes;
    pr.l.push_back(L);
  }

  void norms(Proc& pr, std::vector<NormDesc> v) {
    int64_t t = 0;
    size_t part = 0;
    for (auto& x : v) {
      x.nchunks = int(cdiv(x.n, kNormChunk));
      x.tile_begin = t;
      x.partial = (double*)(uintptr_t)(part * 8);
      t += x.nchunks;
      part += x.nchunks;
    }
    scratch_need = std::max(scratch_need, part);
    for (auto& x : v) x.partial = (double*)(P->scratch.p + (uintptr_t)x.partial);
    size_t off = blob.add(v);
    Launch L;
    L.kind = K_NORM; L.cat = CAT_NORM; L.off = off; L.count = int(v.size()); L.tiles = t;
    pr.l.push_back(L);
    Launch F;
    F.kind = K_NORM_FIN; F.cat = CAT_NORM; F.off = off; F.count = int(v.size());
    pr.l.push_back(F);
  }

  // bf16 plans' reduce W = A^T B (A: the bf16 V copies, B: fp32 rows of X,
  // W fp64): tiles of kReduceRows x kReduceCols, the K range cut into chunks
  // of a multiple of 32 rows until the batch has about min_wg workgroups;
  // chunk partials are added in order by the sum pass.  Not tunable (the
  // tile engine does not read bf16 operands).
  void reduce_bf16(Proc& pr, Cat cat, const std::vector<G>& gs, int64_t min_wg) {
    double work = 0;
    for (const G& g : gs)
      if (g.m > 0 && g.n > 0) work += double(cdiv(g.m, kReduceRows) * cdiv(g.n, kReduceCols)) * std::max(g.k, 1);
    const int64_t chunk = cdiv(std::max<int64_t>(2048, int64_t(work / double(min_wg))), 32) * 32;
    std::vector<GemmDesc> d;
    std::vector<SumDesc> sums;
    size_t part = 0;
    double flops = 0, bytes = 0;
    int64_t tiles = 0;
    for (const G& g : gs) {
      if (g.m <= 0 || g.n <= 0 || g.k <= 0) continue;
      GemmDesc x{};
      x.A = g.A; x.B = g.B; x.Cout = g.Cout; x.lda = g.lda; x.ldb = g.ldb; x.ldcout = g.ldcout;
      x.m = g.m; x.n = g.n; x.k = g.k; x.alpha = g.alpha; x.beta = 0.0;
      int nsplit = g.k >= 2 * chunk ? int(cdiv(g.k, chunk)) : 1;
      x.kchunk = int(cdiv(cdiv(g.k, nsplit), 32) * 32);
      x.nsplit = nsplit = int(cdiv(g.k, x.kchunk));
      if (nsplit > 1) {
        x.Cout = (void*)(uintptr_t)(part * 8);   // patched to the scratch base below
        x.ldcout = g.m;
        x.split_stride = int64_t(g.m) * g.n;
        SumDesc sd{};
        sd.P = (const void*)(uintptr_t)(part * 8);
        sd.stride = x.split_stride;
        sd.out = g.Cout;
        sd.ldo = g.ldcout;
        sd.m = g.m; sd.n = g.n; sd.nsplit = nsplit; sd.ldp = g.m;
        sums.push_back(sd);
        part += size_t(nsplit) * g.m * g.n;
      }
      x.tiles_m = int(cdiv(g.m, kReduceRows));
      x.tiles_n = int(cdiv(g.n, kReduceCols));
      x.tile_begin = tiles;
      tiles += int64_t(x.tiles_m) * x.tiles_n * nsplit;
      d.push_back(x);
      flops += 2.0 * g.m * g.n * g.k;
      bytes += double(g.k) * g.m * 2 + double(g.k) * g.n * 4 + double(g.m) * g.n * 8;
    }
    if (d.empty()) return;
    if (part) {
      scratch_need = std::max(scratch_need, part);
      for (auto& x : d)
        if (x.nsplit > 1) x.Cout = (void*)(P->scratch.p + (uintptr_t)x.Cout);
      for (auto& s : sums) s.P = (const void*)(P->scratch.p + (uintptr_t)s.P);
    }
    Launch L;
    L.kind = K_GEMM; L.cat = cat;
    L.t[0] = BF16; L.t[1] = F32; L.t[2] = F64; L.t[3] = BF16;
    L.trans = true; L.role = 1; L.tile = 4; L.alt[0] = L.alt[1] = 4;
    L.count = int(d.size()); L.flops = flops; L.bytes = bytes;
    L.off = L.off_t[4] = blob.add(d);
    L.tiles = L.tiles_t[4] = tiles;
    pr.l.push_back(L);
    sum(pr, cat == CAT_REDUCE ? CAT_REDUCE_SUM : cat, sums);
  }

  // FP64 products with few columns (the residual and matvec sweeps), given as
  // streaming descriptors: A B products at most kStreamCols wide on the
  // streaming kernel (A read once, no LDS), V^T x products and wider ones on
  // the tile engine's 64 x 32 LDS tiles (two launches when there is a second
  // segment).  Measured at N = 2^20, 1024-row leaves, 32 columns: the rows
  // product K x + U Z runs at 0.68 of HBM streaming (vector loads) against 0.58
  // on the engine, while V^T x streams at 0.35-0.38 (a wave's 64 columns of V,
  // 8 MB apart, each read in 128-256 byte pieces) against 0.55-0.59 on the
  // engine's LDS tiles; at 8 columns (N = 262,144, rank 100) 0.34 streaming
  // against 0.55 on the engine
  void few_cols(Proc& pr, Cat cat, bool trans, const std::vector<StreamDesc>& v) {
    int nmax = 0;
    for (auto& x : v) nmax = std::max(nmax, x.n);
    if (!trans && nmax <= kStreamCols) {
      stream(pr, cat, v);
      return;
    }
    std::vector<G> g1, g2;   // op(A1) B1 (+ beta Cin), then += A2 B2 in a second launch
    for (auto& x : v) {
      G a{};
      a.A = x.A1; a.lda = x.lda1; a.B = x.B1; a.ldb = x.ldb1; a.Cin = x.Cin; a.ldcin = x.ldcin;
      a.Cout = x.Cout; a.ldcout = x.ldcout; a.m = x.m; a.n = x.n; a.k = x.k1; a.alpha = x.alpha; a.beta = x.beta;
      g1.push_back(a);
      if (x.k2 > 0) {
        G b = a;
        b.A = x.A2; b.lda = x.lda2; b.B = x.B2; b.ldb = x.ldb2; b.Cin = x.Cout; b.ldcin = x.ldcout;
        b.k = x.k2; b.beta = 1.0;
        g2.push_back(b);
      }
    }
    gemm(pr, cat, F64, F64, F64, F64, trans, 4, g1);
    gemm(pr, cat, F64, F64, F64, F64, false, 4, g2);
  }

  // FP64 streaming products C = alpha (A1 B1 + A2 B2) + beta Cin, tiles of
  // kStreamRows rows x stream_tile_cols(nmax) columns
  void stream(Proc& pr, Cat cat, std::vector<StreamDesc> v) {
    int nmax = 0;
    for (auto& x : v) nmax = std::max(nmax, x.n);
    const int nc = stream_tile_cols(nmax);
    int64_t t = 0;
    double flops = 0, bytes = 0;
    std::vector<StreamDesc> d;
    // 32-byte vector loads need aligned operands and whole row quads
    auto al = [](const void* ptr, int64_t ld) { return (uintptr_t)ptr % 32 == 0 && ld % 4 == 0; };
    bool vec = true;
    for (auto& x : v) {
      if (x.m <= 0 || x.n <= 0 || x.k1 + x.k2 <= 0) continue;
      if (!al(x.A1, x.lda1) || !al(x.B1, x.ldb1) || x.m % 4 != 0 ||
          (x.k2 > 0 && (!al(x.A2, x.lda2) || !al(x.B2, x.ldb2))))
        vec = false;
    }
    for (auto& x : v) {
      if (x.m <= 0 || x.n <= 0 || x.k1 + x.k2 <= 0) continue;
      x.tiles_m = int(cdiv(x.m, kStreamRows));
      x.tiles_n = int(cdiv(x.n, nc));
      x.tile_begin = t;
      t += int64_t(x.tiles_m) * x.tiles_n;
      flops += 2.0 * x.m * x.n * (double(x.k1) + x.k2);
      bytes += 8.0 * (double(x.m) * (double(x.k1) + x.k2) + (double(x.k1) + x.k2) * x.n +
                      double(x.m) * x.n * (x.beta != 0.0 ? 2 : 1));
      d.push_back(x);
    }
    if (d.empty()) return;
    Launch L;
    L.kind = K_STREAM; L.cat = cat; L.nmax = nmax; L.mode = vec;
    L.off = blob.add(d); L.count = int(d.size()); L.tiles = t; L.flops = flops; L.bytes = bytes;
    pr.l.push_back(L);
  }

  // refinement convergence check over every system's residual and rhs norms
  void check(Proc& pr, int mode) {
    std::vector<CheckDesc> v;
    for (SysState& S : P->sys) {
      CheckDesc c{};
      c.rnorm = (const double*)(P->ws.p + S.rnorm);
      c.bnorm = (const double*)(P->ws.p + S.bnorm);
      c.c = S.tree.nrhs;
      v.push_back(c);
    }
    Launch L;
    L.kind = K_CHECK; L.cat = CAT_NORM; L.off = blob.add(v); L.count = int(v.size()); L.mode = mode;
    pr.l.push_back(L);
  }
};

template <class T> T* at(hodlr_plan_s* P, size_t byte_off) { return (T*)(P->ws.p + byte_off); }

// Workspace layout: one allocation, every region 256-byte aligned.
struct Layout {
  size_t top = 0;
  size_t take(size_t bytes) {
    size_t o = (top + 255) & ~size_t(255);
    top = o + bytes;
    return o;
  }
};

int layout_plan(hodlr_plan_s* P, Layout& L) {
  const size_t we = esize(P->wt);
  int node_count = 0;
  for (SysState& S : P->sys) {
    const Tree& t = S.tree;
    const int64_t n = t.n, c = t.nrhs;
    S.ldw = n;
    S.xws = L.take(size_t(n) * t.xcols() * we);
    S.yws = L.take(size_t(n) * c * we);
    S.rbuf = L.take(size_t(n) * c * 8);
    S.mvx = L.take(size_t(n) * c * 8);
    if (we != 8 && P->mt != BF16) S.vws = L.take(size_t(n) * std::max<int64_t>(t.vcols(), 1) * we);
    S.vb.assign(t.nodes.size(), 0);
    if (P->mt == BF16)
      for (size_t i = 1; i < t.nodes.size(); ++i)
        S.vb[i] = L.take(size_t(cdiv(t.nodes[i].nrow, 32) * 32) * size_t(cdiv(t.rank(t.nodes[i].depth), 16) * 16) * 2);
    S.lpiv = L.take(size_t(n) * 4);
    S.lperm = L.take(size_t(n) * 4);
    size_t lu_elems = 0, z_elems = 0, k_elems = 0;
    size_t dinv_elems = 0;
    S.lu_off.resize(t.leaves.size());
    S.dinv_off.resize(t.leaves.size());
    S.z_off.resize(t.leaves.size());
    S.kofs.resize(t.leaves.size());
    for (size_t j = 0; j < t.leaves.size(); ++j) {
      const TreeNode& x = t.nodes[t.leaves[j]];
      S.lu_off[j] = lu_elems;
      lu_elems += size_t(x.nrow) * x.nrow;
      S.dinv_off[j] = dinv_elems;
      dinv_elems += size_t(cdiv(x.nrow, 64)) * 2 * 64 * 64;
      S.z_off[j] = z_elems;
      z_elems += size_t(t.rsum[x.depth]) * c;
      S.kofs[j] = S.s.klayout == HODLR_K_STACKED ? size_t(x.row_beg) : k_elems;
      k_elems += size_t(x.nrow) * x.nrow;
    }
    S.lu = L.take(lu_elems * we);
    S.ldinv = L.take(dinv_elems * we);
    if (P->mt == BF16) {
      size_t lut_elems = 0;
      S.lut_off.resize(t.leaves.size());
      for (size_t j = 0; j < t.leaves.size(); ++j) {
        const int64_t m = t.nodes[t.leaves[j]].nrow;
        S.lut_off[j] = lut_elems;
        lut_elems += size_t(cdiv(m, 16) * 16) * size_t(cdiv(m, 32) * 32);
      }
      S.lut = L.take(lut_elems * 2);
    }
    S.zbuf = L.take(std::max<size_t>(z_elems, 1) * 8);
    const size_t nn = t.nodes.size();
    S.leaf_lo.assign(nn, int(t.leaves.size()));
    S.leaf_hi.assign(nn, 0);
    S.pld = 0;
    for (size_t j = 0; j < t.leaves.size(); ++j) {
      S.pld = std::max<int64_t>(S.pld, t.rsum[t.nodes[t.leaves[j]].depth]);
      for (int a = t.leaves[j];; a = t.nodes[a].parent) {
        S.leaf_lo[a] = std::min(S.leaf_lo[a], int(j));
        S.leaf_hi[a] = std::max(S.leaf_hi[a], int(j) + 1);
        if (a == 0) break;
      }
    }
    S.pmv = L.take(std::max<size_t>(t.leaves.size() * size_t(S.pld) * c, 1) * 8);
    S.vtd.assign(nn, 0); S.vtu.assign(nn, 0); S.wmv.assign(nn, 0);
    S.sbuf.assign(nn, 0); S.eta.assign(nn, 0); S.spiv.assign(nn, 0); S.sperm.assign(nn, 0); S.sdinv.assign(nn, 0);
    S.etaf.assign(nn, 0);
    S.sinfo.assign(nn, -1);
    for (size_t i = 1; i < nn; ++i) {
      const TreeNode& x = t.nodes[i];
      const int r = t.rank(x.depth);
      const int64_t nd = t.col_beg(x.depth);
      // V^T d then V^T u, adjacent (ld r): the reduce writes both as one
      // r x (nd + r) product (nd = c + R_{depth-1} >= c)
      S.vtd[i] = L.take(size_t(r) * (nd + r) * 8);
      S.vtu[i] = S.vtd[i] + size_t(r) * nd * 8;
      S.wmv[i] = L.take(size_t(r) * c * 8);
    }
    for (size_t i = 0; i < nn; ++i) {
      const TreeNode& x = t.nodes[i];
      if (x.is_leaf()) continue;
      const int r = t.rank(x.depth + 1);
      const int64_t nd = t.col_beg(x.depth + 1);
      S.sbuf[i] = L.take(size_t(r) * r * 8);
      S.eta[i] = L.take(size_t(r) * std::max<int64_t>(nd, c) * 8);
      if (we != 8) {
        // fp32 copies of both etas (ld r rounded up to 8) or, in bf16 plans, their bf16 MFMA fragments
        const int64_t ncol = std::max<int64_t>(nd, c);
        const size_t f32 = size_t(cdiv(r, 8) * 8) * ncol * we, frag = size_t(cdiv(r, 32) * 32) * cdiv(ncol, 16) * 16 * 2;
        S.etaf[i] = L.take(2 * std::max(f32, frag));
      }
      S.spiv[i] = L.take(size_t(r) * 4);
      S.sperm[i] = L.take(size_t(r) * 4);
      S.sdinv[i] = L.take(size_t(cdiv(r, 64)) * 2 * 64 * 64 * 8);
      S.sinfo[i] = node_count++;
    }
    S.bnorm = L.take(size_t(c) * 8);
    S.rnorm = L.take(size_t(c) * 8);
  }
  P->nnodes = node_count;
  return 0;
}

// Matvec reduce: W_i = V_i^T x_i for every non-root node i (V_i: the rows of
// node i in the column block of its depth), for the HODLR product of the
// residual.  One streaming product per leaf reads each V row block once,
// P_j = V(rows of leaf j, 0 : R_D)^T x_j, and W_i is then the sum, in leaf
// order, of the depth block of P_j over the leaves j under node i.
void matvec_reduce(hodlr_plan_s* P, SysState& S, const double* x, int64_t ldx, std::vector<StreamDesc>& prod,
                   std::vector<SumDesc>& sums) {
  const Tree& t = S.tree;
  const int64_t c = t.nrhs;
  double* Pb = at<double>(P, S.pmv);
  const int64_t stride = S.pld * c;
  for (size_t j = 0; j < t.leaves.size(); ++j) {
    const TreeNode& nd = t.nodes[t.leaves[j]];
    const int64_t Rd = t.rsum[nd.depth];
    if (Rd == 0) continue;
    StreamDesc g{};
    g.A1 = S.s.V + nd.row_beg; g.lda1 = S.s.ldv;
    g.B1 = x + nd.row_beg; g.ldb1 = ldx;
    g.Cout = Pb + int64_t(j) * stride; g.ldcout = S.pld;
    g.m = int(Rd); g.n = int(c); g.k1 = int(nd.nrow); g.alpha = 1.0; g.beta = 0.0;
    prod.push_back(g);
  }
  for (size_t i = 1; i < t.nodes.size(); ++i) {
    const TreeNode& nd = t.nodes[i];
    const int r = t.rank(nd.depth);
    SumDesc sd{};
    sd.P = Pb + int64_t(S.leaf_lo[i]) * stride + t.rsum[nd.depth - 1];
    sd.stride = stride; sd.ldp = int(S.pld);
    sd.out = at<double>(P, S.wmv[i]); sd.ldo = r;
    sd.m = r; sd.n = int(c); sd.nsplit = S.leaf_hi[i] - S.leaf_lo[i];
    sums.push_back(sd);
  }
}

// HODLR product into the plan's rbuf, for every system:
//   from_solution: rbuf = b + sign * A x with x the caller's solution array and
//                  b = X[:, 0:c] (sign = -1: the residual b - A x);
//   otherwise:     rbuf = sign * A mvx (the staged input of hodlr_matvec).
// A x = blockdiag(K) x + sum over nodes of U_child (V_sibling^T x_sibling):
// per leaf j, the rows of A x are [K_j | U_j] [x_j ; Z_j] with Z_j the
// gathered V^T x blocks of its ancestors' siblings -- one two-segment
// streaming product that reads K_j and U_j once -- all in FP64.
void hodlr_product(hodlr_plan_s* P, Builder& B, Proc& pr, bool from_solution, double sign) {
  std::vector<StreamDesc> red, rows;
  std::vector<SumDesc> wsum;
  std::vector<CopyDesc> gz;
  for (SysState& S : P->sys) {
    const Tree& t = S.tree;
    const int64_t c = t.nrhs;
    double* R = at<double>(P, S.rbuf);
    double* Z = at<double>(P, S.zbuf);
    const double* x = from_solution ? S.s.x : at<double>(P, S.mvx);
    const int64_t ldx = from_solution ? S.s.ldxs : S.ldw;
    matvec_reduce(P, S, x, ldx, red, wsum);
    for (size_t j = 0; j < t.leaves.size(); ++j) {
      const int id = t.leaves[j];
      const TreeNode& nd = t.nodes[id];
      const int64_t Rd = t.rsum[nd.depth];
      StreamDesc g{};
      g.A1 = S.s.K + S.kofs[j]; g.lda1 = S.s.klayout == HODLR_K_STACKED ? S.s.ldk : nd.nrow;
      g.B1 = x + nd.row_beg; g.ldb1 = ldx; g.k1 = int(nd.nrow);
      if (Rd > 0) {
        g.A2 = S.s.X + nd.row_beg + c * S.s.ldx; g.lda2 = S.s.ldx;
        g.B2 = Z + S.z_off[j]; g.ldb2 = Rd; g.k2 = int(Rd);
      }
      g.Cin = from_solution ? S.s.X + nd.row_beg : nullptr; g.ldcin = S.s.ldx;
      g.Cout = R + nd.row_beg; g.ldcout = S.ldw;
      g.m = int(nd.nrow); g.n = int(c); g.alpha = sign; g.beta = from_solution ? 1.0 : 0.0;
      rows.push_back(g);
      if (Rd == 0) continue;
      // Z_j rows of depth l: w of the sibling of the depth-l ancestor
      int a = id;
      while (a != 0) {
        const TreeNode& an = t.nodes[a];
        const TreeNode& par = t.nodes[an.parent];
        const int sib = par.child[0] == a ? par.child[1] : par.child[0];
        const int r = t.rank(an.depth);
        CopyDesc cd{};
        cd.src = at<double>(P, S.wmv[sib]); cd.lds = r;
        cd.dst = Z + S.z_off[j] + t.rsum[an.depth - 1]; cd.ldd = Rd;
        cd.rows = r; cd.cols = int(c); cd.alpha = 1.0;
        gz.push_back(cd);
        a = an.parent;
      }
    }
  }
  B.few_cols(pr, CAT_MATVEC_VTX, true, red);
  B.sum(pr, CAT_MATVEC, wsum);
  B.copy(pr, CAT_MATVEC, F64, F64, gz);
  B.few_cols(pr, CAT_MATVEC_ROWS, false, rows);
}

void build_all(hodlr_plan_s* P, Builder& B) {
  const DT wt = P->wt;
  const size_t we = esize(wt);
  const int H = P->height;
  P->reduce.assign(H + 1, Proc());
  P->node.assign(H + 1, Proc());
  P->bcast.assign(H + 1, Proc());
  P->reduce_s.assign(H + 1, Proc());
  P->node_s.assign(H + 1, Proc());
  P->bcast_s.assign(H + 1, Proc());
  int leaf_base = 0;

  // ---------------------------------------------------------------- leaves
  {
    std::vector<CopyDesc> kc;
    std::vector<LuDesc> lus;
    std::vector<TrsmDesc> tr, trs;
    std::vector<TransDesc> tt;
    for (SysState& S : P->sys) {
      const Tree& t = S.tree;
      for (size_t j = 0; j < t.leaves.size(); ++j) {
        const TreeNode& x = t.nodes[t.leaves[j]];
        const int m = int(x.nrow);
        CopyDesc c{};
        c.src = S.s.K + S.kofs[j];
        c.lds = S.s.klayout == HODLR_K_STACKED ? S.s.ldk : x.nrow;
        c.dst = P->ws.p + S.lu + S.lu_off[j] * we;
        c.ldd = m;
        c.rows = c.cols = m;
        c.alpha = 1.0;
        kc.push_back(c);
        LuDesc l{};
        l.A = P->ws.p + S.lu + S.lu_off[j] * we;
        l.lda = m;
        l.m = m;
        l.ipiv = at<int32_t>(P, S.lpiv) + x.row_beg;
        l.perm = at<int32_t>(P, S.lperm) + x.row_beg;
        l.info = P->leaf_info + leaf_base + j;
        l.dinv = P->ws.p + S.ldinv + S.dinv_off[j] * we;
        lus.push_back(l);
        TrsmDesc d{};
        d.LU = l.A; d.ldlu = m; d.perm = l.perm; d.dinv = l.dinv;
        if (P->mt == BF16) {
          d.LUt = P->ws.p + S.lut + S.lut_off[j] * 2;
          d.ldt = cdiv(m, 32) * 32;
          TransDesc x{};
          x.src = l.A; x.lds = m; x.dst = (void*)d.LUt; x.ldd = d.ldt; x.m = m; x.tiles_k = int(cdiv(d.ldt, 64));
          tt.push_back(x);
        }
        d.Xin = S.s.X + x.row_beg; d.ldin = S.s.ldx;
        d.Xout = P->ws.p + S.xws + size_t(x.row_beg) * we; d.ldout = S.ldw;
        d.m = m; d.ncols = int(t.leaf_nrhs(t.leaves[j]));
        tr.push_back(d);
        TrsmDesc e = d;
        e.Xin = at<double>(P, S.rbuf) + x.row_beg; e.ldin = S.ldw;
        e.Xout = P->ws.p + S.yws + size_t(x.row_beg) * we; e.ldout = S.ldw;
        e.ncols = t.nrhs;
        trs.push_back(e);
      }
      leaf_base += int(t.leaves.size());
    }
    if (P->mt == BF16) {
      // bf16 plans: every node's V block rounded to bf16 once per factor pass
      // (through fp32, as the products would round it), stored as the MFMA
      // A fragments of V^T that the reduce loads directly
      std::vector<FragDesc> vc;
      for (SysState& S : P->sys) {
        const Tree& t = S.tree;
        for (size_t i = 1; i < t.nodes.size(); ++i) {
          const TreeNode& x = t.nodes[i];
          FragDesc f{};
          f.src = S.s.V + x.row_beg + t.rsum[x.depth - 1] * S.s.ldv; f.lds = S.s.ldv;
          f.dst = P->ws.p + S.vb[i];
          f.rows = int(x.nrow); f.cols = t.rank(x.depth); f.ngroups = int(cdiv(f.cols, 16));
          vc.push_back(f);
        }
      }
      B.frag(P->leaf, CAT_LEAF_COPY, vc);
    } else if (wt != F64) {
      // V is read by every column tile of the reduce: round it to FP32 once
      std::vector<CopyDesc> vc;
      for (SysState& S : P->sys) {
        CopyDesc c{};
        c.src = S.s.V; c.lds = S.s.ldv; c.dst = P->ws.p + S.vws; c.ldd = S.ldw;
        c.rows = int(S.tree.n); c.cols = int(S.tree.vcols()); c.alpha = 1.0;
        vc.push_back(c);
      }
      B.copy(P->leaf, CAT_LEAF_COPY, F64, wt, vc);
    }
    B.copy(P->leaf, CAT_LEAF_COPY, F64, wt, kc);
    B.lu(P->leaf, CAT_LU_PANEL, CAT_LU_ROW, CAT_LU_UPDATE, CAT_LU_MISC, wt, lus, P->mt);
    B.transpose(P->leaf, CAT_LU_MISC, tt);
    B.trsm(P->leaf, CAT_LEAF_TRSM, wt, F64, wt, tr, P->mt);
    B.trsm(P->solve_only_leaf, CAT_SOLVE_TRSM, wt, F64, wt, trs, P->mt);
  }

  // ---------------------------------------------------------------- heights
  for (int h = 1; h <= H; ++h) {
    for (int mode = 0; mode < 2; ++mode) {   // 0: fused factor+solve, 1: solve-only
      std::vector<Builder::G> red, ng1, ng2, bc;
      std::vector<CopyDesc> sform, eround;
      std::vector<LuDesc> slu;
      std::vector<TrsmDesc> str;
      std::vector<FragDesc> efrag;
      // bf16 plans hand the factor's broadcast its etas as bf16 fragments when
      // every node rank of the height fits the row-panel kernel (k <= kPanelKMax;
      // measured 14.0 -> 12.1 ms at N = 2^20, c = 32, rank 128).  The
      // correction solves' broadcasts (c columns) keep fp32 etas and the
      // tunable choice between the row panel and the 64 x 32 engine tiles.
      bool bfrag = P->mt == BF16 && mode == 0;
      for (SysState& S : P->sys)
        for (const TreeNode& Pn : S.tree.nodes)
          if (!Pn.is_leaf() && Pn.height == h && S.tree.rank(Pn.depth + 1) > kPanelKMax) bfrag = false;
      for (SysState& S : P->sys) {
        const Tree& t = S.tree;
        const int64_t c = t.nrhs;
        char* X = P->ws.p + (mode == 0 ? S.xws : S.yws);
        char* F = P->ws.p + S.xws;   // factor columns
        for (size_t pi = 0; pi < t.nodes.size(); ++pi) {
          const TreeNode& Pn = t.nodes[pi];
          if (Pn.is_leaf() || Pn.height != h) continue;
          const int d = Pn.depth + 1;
          const int r = t.rank(d);
          const int64_t ucol = t.col_beg(d), vcol = t.rsum[d - 1];
          const int nd = int(mode == 0 ? ucol : c);
          const int ch[2] = {Pn.child[0], Pn.child[1]};
          double* vtd[2] = {at<double>(P, S.vtd[ch[0]]), at<double>(P, S.vtd[ch[1]])};
          double* vtu[2] = {at<double>(P, S.vtu[ch[0]]), at<double>(P, S.vtu[ch[1]])};
          double* eta0 = at<double>(P, S.eta[pi]);
          double* Sm = at<double>(P, S.sbuf[pi]);
          // W = V^T [d | u] (mode 0; V^T d = vtd and V^T u = vtu adjacent,
          // X's d and u columns adjacent: one product) or V^T y (mode 1)
          for (int k = 0; k < 2; ++k) {
            const TreeNode& C = t.nodes[ch[k]];
            Builder::G g{};
            if (P->mt == BF16) {
              g.A = P->ws.p + S.vb[ch[k]]; g.lda = cdiv(r, 16);   // fragment column groups
            } else if (wt == F64) {
              g.A = S.s.V + C.row_beg + vcol * S.s.ldv; g.lda = S.s.ldv;
            } else {
              g.A = P->ws.p + S.vws + (size_t(C.row_beg) + size_t(vcol) * S.ldw) * we; g.lda = S.ldw;
            }
            g.B = X + size_t(C.row_beg) * we; g.ldb = S.ldw;
            g.Cout = vtd[k]; g.ldcout = r; g.Cin = nullptr; g.ldcin = r;
            g.m = r; g.n = mode == 0 ? nd + r : nd; g.k = int(C.nrow); g.alpha = 1.0; g.beta = 0.0;
            red.push_back(g);
          }
          // node: VTd1 -= VTu1 VTd0
          Builder::G a{};
          a.A = vtu[1]; a.lda = r; a.B = vtd[0]; a.ldb = r; a.Cin = vtd[1]; a.ldcin = r; a.Cout = vtd[1]; a.ldcout = r;
          a.m = r; a.n = nd; a.k = r; a.alpha = -1.0; a.beta = 1.0;
          ng1.push_back(a);
          if (mode == 0) {
            Builder::G s{};
            s.A = vtu[1]; s.lda = r; s.B = vtu[0]; s.ldb = r; s.Cout = Sm; s.ldcout = r; s.Cin = nullptr; s.ldcin = r;
            s.m = r; s.n = r; s.k = r; s.alpha = 1.0; s.beta = 0.0;
            ng1.push_back(s);
            CopyDesc cd{};
            cd.src = Sm; cd.lds = r; cd.dst = Sm; cd.ldd = r; cd.rows = cd.cols = r;
            cd.alpha = -1.0; cd.beta = 0.0; cd.diag = 1.0;
            sform.push_back(cd);
            LuDesc l{};
            l.A = Sm; l.lda = r; l.m = r;
            l.ipiv = at<int32_t>(P, S.spiv[pi]);
            l.perm = at<int32_t>(P, S.sperm[pi]);
            l.info = P->node_info + S.sinfo[pi];
            l.dinv = P->ws.p + S.sdinv[pi];
            slu.push_back(l);
          }
          TrsmDesc td{};
          td.LU = Sm; td.ldlu = r; td.perm = at<int32_t>(P, S.sperm[pi]); td.dinv = P->ws.p + S.sdinv[pi];
          td.Xin = vtd[1]; td.ldin = r; td.Xout = eta0; td.ldout = r;
          td.m = r; td.ncols = nd;
          str.push_back(td);
          Builder::G b{};
          b.A = vtu[0]; b.lda = r; b.B = eta0; b.ldb = r; b.Cin = vtd[0]; b.ldcin = r; b.Cout = vtd[0]; b.ldcout = r;
          b.m = r; b.n = nd; b.k = r; b.alpha = -1.0; b.beta = 1.0;
          ng2.push_back(b);
          // broadcast: d0 -= u0 eta0, d1 -= u1 eta1 (eta1 in the VTd0 slot); an
          // FP32 plan first rounds both etas to FP32 once, so the broadcast
          // tiles stream 4-byte operands
          const void* etas[2] = {eta0, vtd[0]};
          // (leading dimension r rounded up to 8, rows past r zero: the plan's
          // workspace is cleared at creation and those rows are never written)
          const int rp = int(cdiv(r, 8) * 8);
          int ldeta = r;
          if (bfrag) {
            // bf16 plans: both etas as the bf16 MFMA fragments the row-panel
            // broadcast loads directly (the rounding its products would do)
            char* ef = P->ws.p + S.etaf[pi];
            const size_t fbytes = size_t(cdiv(r, 32) * 32) * cdiv(nd, 16) * 16 * 2;
            for (int k = 0; k < 2; ++k) {
              FragDesc f{};
              f.src = (const double*)etas[k]; f.lds = r; f.dst = ef + size_t(k) * fbytes;
              f.rows = r; f.cols = nd; f.ngroups = int(cdiv(nd, 16));
              efrag.push_back(f);
              etas[k] = f.dst;
            }
            ldeta = int(cdiv(nd, 16));
          } else if (wt != F64) {
            char* ef = P->ws.p + S.etaf[pi];
            for (int k = 0; k < 2; ++k) {
              CopyDesc cd{};
              cd.src = etas[k]; cd.lds = r; cd.dst = ef + size_t(k) * rp * nd * we; cd.ldd = rp;
              cd.rows = r; cd.cols = nd; cd.alpha = 1.0;
              eround.push_back(cd);
              etas[k] = cd.dst;
            }
            ldeta = rp;
          }
          for (int k = 0; k < 2; ++k) {
            const TreeNode& C = t.nodes[ch[k]];
            Builder::G g{};
            g.A = F + (size_t(C.row_beg) + size_t(ucol) * S.ldw) * we; g.lda = S.ldw;
            g.B = etas[k]; g.ldb = ldeta;
            g.Cout = X + size_t(C.row_beg) * we; g.Cin = X + size_t(C.row_beg) * we; g.ldcin = g.ldcout = S.ldw;
            g.m = int(C.nrow); g.n = nd; g.k = r; g.alpha = -1.0; g.beta = 1.0;
            bc.push_back(g);
          }
        }
      }
      Proc& pr = mode == 0 ? P->reduce[h] : P->reduce_s[h];
      Proc& pn = mode == 0 ? P->node[h] : P->node_s[h];
      Proc& pb = mode == 0 ? P->bcast[h] : P->bcast_s[h];
      const Cat cr = mode == 0 ? CAT_REDUCE : CAT_SOLVE_REDUCE;
      const Cat cg = mode == 0 ? CAT_NODE_GEMM : CAT_SOLVE_NODE;
      const Cat ct = mode == 0 ? CAT_NODE_TRSM : CAT_SOLVE_NODE;
      const Cat cb = mode == 0 ? CAT_BCAST : CAT_SOLVE_BCAST;
      if (P->mt == BF16) B.reduce_bf16(pr, cr, red, 2048);
      else B.gemm(pr, cr, wt, wt, F64, P->mt, true, 1, red, 2048);
      B.gemm(pn, cg, F64, F64, F64, F64, false, 0, ng1);
      if (mode == 0) {
        B.copy(pn, CAT_NODE_LU, F64, F64, sform);
        B.lu(pn, CAT_NODE_LU, CAT_NODE_LU, CAT_NODE_LU, CAT_NODE_LU, F64, slu);
      }
      B.trsm(pn, ct, F64, F64, F64, str);
      B.gemm(pn, cg, F64, F64, F64, F64, false, 0, ng2);
      B.copy(pn, cg, F64, wt, eround);
      B.frag(pn, cg, efrag);
      if (bfrag) B.bcast_frag(pb, cb, bc);
      else B.gemm(pb, cb, wt, wt == F64 ? F64 : wt, wt, P->mt, false, 2, bc);
    }
  }

  // ---------------------------------------------------------------- extract / update
  {
    std::vector<CopyDesc> ex, up, in;
    for (SysState& S : P->sys) {
      CopyDesc c{};
      c.src = P->ws.p + S.xws; c.lds = S.ldw; c.dst = S.s.x; c.ldd = S.s.ldxs;
      c.rows = int(S.tree.n); c.cols = S.tree.nrhs; c.alpha = 1.0;
      ex.push_back(c);
      CopyDesc u = c;
      u.src = P->ws.p + S.yws; u.beta = 1.0;
      up.push_back(u);
    }
    B.copy(P->extract, CAT_COPY, wt, F64, ex);
    B.copy(P->update, CAT_COPY, wt, F64, up);
  }

  // ---------------------------------------------------------------- residual r = b - A x, ||b||, ||r||
  {
    std::vector<NormDesc> nr;
    for (SysState& S : P->sys) {
      const int64_t c = S.tree.nrhs;
      for (int64_t j = 0; j < c; ++j) {
        NormDesc d{};
        d.x = at<double>(P, S.rbuf) + j * S.ldw; d.n = S.tree.n; d.out = at<double>(P, S.rnorm) + j;
        nr.push_back(d);
        // ||b_j|| is taken at every residual: the caller may change b between solves
        NormDesc e{};
        e.x = S.s.X + j * S.s.ldx; e.n = S.tree.n; e.out = at<double>(P, S.bnorm) + j;
        nr.push_back(e);
      }
    }
    hodlr_product(P, B, P->residual, /*from_solution=*/true, -1.0);
    B.norms(P->residual, nr);
  }
  // ---------------------------------------------------------------- hodlr_matvec: rbuf = A mvx
  hodlr_product(P, B, P->matvec, /*from_solution=*/false, 1.0);
  // ---------------------------------------------------------------- correction solve result: rbuf = yws
  {
    std::vector<CopyDesc> yo;
    for (SysState& S : P->sys) {
      CopyDesc c{};
      c.src = P->ws.p + S.yws; c.lds = S.ldw; c.dst = at<double>(P, S.rbuf); c.ldd = S.ldw;
      c.rows = int(S.tree.n); c.cols = S.tree.nrhs; c.alpha = 1.0;
      yo.push_back(c);
    }
    B.copy(P->solve_out, CAT_COPY, wt, F64, yo);
  }
  // ---------------------------------------------------------------- device-driven refinement
  // first: the residual of the fused pass's solution and its check; sweep:
  // the correction solve on that residual (already in rbuf), x += y, the new
  // residual and its check, every launch gated by the check before it
  P->first = P->residual;
  B.check(P->first, 0);
  P->sweep = Proc();
  P->sweep.append(P->solve_only_leaf);
  for (int h = 1; h <= P->height; ++h) {
    P->sweep.append(P->reduce_s[h]);
    P->sweep.append(P->node_s[h]);
    P->sweep.append(P->bcast_s[h]);
  }
  P->sweep.append(P->update);
  P->sweep.append(P->residual);
  B.check(P->sweep, 1);
  for (Launch& L : P->sweep.l) {
    L.gated = true;
    // every kind a sweep holds takes a gate; a launch that ignored it would
    // run in the sweeps the check has switched off
    if (L.kind != K_GEMM && L.kind != K_SUM && L.kind != K_COPY && L.kind != K_TRSM && L.kind != K_NORM &&
        L.kind != K_NORM_FIN && L.kind != K_CHECK && L.kind != K_STREAM && L.kind != K_FRAG)
      B.bad_gate = true;
  }
}

// ------------------------------------------------------------------ execution
hipError_t dispatch(hodlr_plan_s* P, const Launch& L) {
  const void* d = P->descs.p + L.off;
  hipStream_t s = P->stream;
  RefineCtl* ctl = (RefineCtl*)(P->ws.p + P->ctl);
  const int32_t* gate = L.gated ? &ctl->active : nullptr;
  switch (L.kind) {
    case K_GEMM: {
      int tile = L.tile;
      if (L.gid >= 0) {
        if (P->tuning) tile = L.alt[P->tune_cfg];
        else if (L.gid < int(P->tile_choice.size()) && P->tile_choice[L.gid] >= 0) tile = L.alt[P->tile_choice[L.gid]];
      }
      const GemmDesc* gd = (const GemmDesc*)(P->descs.p + L.off_t[tile]);
      if (tile == 3) return launch_panel_gemm(L.t[0], L.t[3], L.mode != 0, gd, L.count, L.tiles_t[3], s, gate);
      if (tile == 4) return launch_reduce_bf16(gd, L.count, L.tiles_t[4], s, gate);
      return launch_gemm(L.t[0], L.t[1], L.t[2], L.t[3], L.trans, L.role, tile, gd, L.count, L.tiles_t[tile], s,
                         gate);
    }
    case K_LU_DINV: return launch_lu_dinv(L.t[0], (const LuDesc*)d, L.count, L.k0 / kLuNB, std::max(1, L.mode), s);
    case K_LU_SMALL: return launch_lu_small(L.t[0], L.mmax, (const LuDesc*)d, L.count, s);
    case K_SUM: return launch_sum((const SumDesc*)d, L.count, L.tiles, s, gate);
    case K_COPY: return launch_copy(L.t[0], L.t[1], (const CopyDesc*)d, L.count, L.tiles, s, gate);
    case K_CHECK: return launch_refine_check((const CheckDesc*)d, L.count, ctl, L.mode, s, gate);
    case K_STREAM: return launch_stream(L.nmax, L.mode != 0, (const StreamDesc*)d, L.count, L.tiles, s, gate);
    case K_FRAG: return launch_frag_bf16(L.mode != 0, (const FragDesc*)d, L.count, L.tiles, s, gate);
    case K_LU_INIT: return launch_lu_init((const LuDesc*)d, L.count, s);
    case K_LU_PANEL: return launch_lu_panel(L.t[0], (const LuDesc*)d, L.count, L.k0, L.rows_max, s);
    case K_LU_PERM: return launch_lu_perm((const LuDesc*)d, L.count, s);
    case K_TRSM: {
      int cfg = L.tile;
      if (L.gid >= 0) {
        if (P->tuning) cfg = P->tune_cfg;
        else if (L.gid < int(P->tile_choice.size()) && P->tile_choice[L.gid] >= 0) cfg = P->tile_choice[L.gid];
      }
      const TrsmDesc* td = (const TrsmDesc*)(P->descs.p + L.off_t[cfg]);
      if (L.t[3] == BF16) return launch_trsm_lds(L.t[1], 2 + cfg, L.mmax, td, L.count, L.tiles_t[cfg], s, gate);
      return launch_trsm(L.t[0], L.t[1], L.t[2], L.t[3], cfg, td, L.count, L.tiles_t[cfg], s, gate);
    }
    case K_TRANS: return launch_transpose_bf16((const TransDesc*)d, L.count, L.tiles, s);
    case K_NORM: return launch_norm((const NormDesc*)d, L.count, L.tiles, s, gate);
    case K_NORM_FIN: return launch_norm_finish((const NormDesc*)d, L.count, s, gate);
  }
  return hipErrorInvalidValue;
}

hipEvent_t take_event(hodlr_plan_s* P) {
  if (!P->pool.empty()) {
    hipEvent_t e = P->pool.back();
    P->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int harvest(hodlr_plan_s* P) {
  if (P->pending.empty()) return HODLR_OK;
  if (hipStreamSynchronize(P->stream) != hipSuccess) return HODLR_ERR_HIP;
  for (auto& q : P->pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, q.a, q.b) != hipSuccess) return HODLR_ERR_HIP;
    P->prof_ms[q.cat] += ms;
    P->prof_n[q.cat] += 1;
    P->prof_flops[q.cat] += q.flops;
    P->prof_bytes[q.cat] += q.bytes;
    P->pool.push_back(q.a);
    P->pool.push_back(q.b);
  }
  P->pending.clear();
  return HODLR_OK;
}

// Drops one captured graph.
void drop_one(hodlr_plan_s* P, hodlr_plan_s::Graph& g) {
  if (g.exec) {
    (void)hipStreamSynchronize(P->stream);
    (void)hipGraphExecDestroy(g.exec);
  }
  g.exec = nullptr;
  g.sweeps = 0;
}

// Drops every captured graph.
void drop_graph(hodlr_plan_s* P) {
  drop_one(P, P->g_factor);
  drop_one(P, P->g_solve);
  drop_one(P, P->g_sweep);
}

int run(hodlr_plan_s* P, const Proc& pr);

// Captures body() on the plan's stream into g (nothing runs).
template <class F> int capture(hodlr_plan_s* P, hodlr_plan_s::Graph& g, F&& body) {
  drop_one(P, g);
  hipGraph_t gr;
  if (hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return HODLR_ERR_HIP;
  int rc = body();
  if (hipStreamEndCapture(P->stream, &gr) != hipSuccess) return HODLR_ERR_HIP;
  if (rc || hipGraphInstantiate(&g.exec, gr, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(gr);
    g.exec = nullptr;
    return rc ? rc : HODLR_ERR_HIP;
  }
  (void)hipGraphDestroy(gr);
  return HODLR_OK;
}

int replay(hodlr_plan_s* P, hodlr_plan_s::Graph& g) {
  return hipGraphLaunch(g.exec, P->stream) == hipSuccess ? HODLR_OK : HODLR_ERR_HIP;
}

int run(hodlr_plan_s* P, const Proc& pr) {
  for (const Launch& L : pr.l) {
    if (P->tuning && (L.kind == K_GEMM || L.kind == K_TRSM) && L.gid >= 0) {
      hipEvent_t a = take_event(P), b = take_event(P);
      if (!a || !b) return HODLR_ERR_HIP;
      if (hipEventRecord(a, P->stream) != hipSuccess) return HODLR_ERR_HIP;
      if (dispatch(P, L) != hipSuccess) return HODLR_ERR_HIP;
      if (hipEventRecord(b, P->stream) != hipSuccess) return HODLR_ERR_HIP;
      P->tune_pending.push_back({L.gid, a, b});
      continue;
    }
    if (P->prof_mask & (1u << L.cat)) {
      hipEvent_t a = take_event(P), b = take_event(P);
      if (!a || !b) return HODLR_ERR_HIP;
      if (hipEventRecord(a, P->stream) != hipSuccess) return HODLR_ERR_HIP;
      if (dispatch(P, L) != hipSuccess) return HODLR_ERR_HIP;
      if (hipEventRecord(b, P->stream) != hipSuccess) return HODLR_ERR_HIP;
      P->pending.push_back({int(L.cat), a, b, L.flops, L.bytes});
    } else if (dispatch(P, L) != hipSuccess) {
      return HODLR_ERR_HIP;
    }
  }
  return HODLR_OK;
}

double proc_flops(const Proc& p) {
  double f = 0;
  for (auto& l : p.l) f += l.flops;
  return f;
}

void fill_info(hodlr_plan_s* P) {
  hodlr_plan_info_t& I = P->info;
  I.nsys = int(P->sys.size());
  I.precision = P->precision;
  I.height = P->height;
  I.nleaves = P->nleaves;
  I.workspace_bytes = int64_t(P->ws.bytes + P->descs.bytes + P->scratch.bytes);
  // algorithmic work (flop and byte model of one fused pass)
  const double we = double(esize(P->wt));
  double fl = 0, fr = 0, fn = 0, fb = 0, bl = 0, br = 0, bb = 0, fm = 0, bm = 0, fs = 0, bs = 0;
  for (SysState& S : P->sys) {
    const Tree& t = S.tree;
    const double c = t.nrhs;
    for (int id : t.leaves) {
      const TreeNode& x = t.nodes[id];
      const double m = double(x.nrow), nr = double(t.leaf_nrhs(id)), Rd = double(t.rsum[x.depth]);
      fl += 2.0 / 3.0 * m * m * m + 2.0 * m * m * nr;
      bl += we * (m * m + 2.0 * m * nr);
      fm += 2.0 * m * m * c + 2.0 * m * Rd * c;
      bm += 8.0 * (m * m + m * Rd);
      fs += 2.0 * m * m * c;
      bs += we * m * m;
    }
    for (const TreeNode& Pn : t.nodes) {
      if (Pn.is_leaf()) continue;
      const int d = Pn.depth + 1;
      const double r = t.rank(d), nd = double(t.col_beg(d));
      for (int k = 0; k < 2; ++k) {
        const double n = double(t.nodes[Pn.child[k]].nrow);
        fr += 2.0 * n * r * (nd + r);
        br += we * n * (nd + r) + 8.0 * n * r;
        fb += 2.0 * n * r * nd;
        bb += we * n * (r + 2.0 * nd);
        fm += 2.0 * 2.0 * n * r * c;
        bm += 8.0 * n * r * 2.0;
        fs += 2.0 * 2.0 * n * r * c;
        bs += 8.0 * n * r + we * n * r;
      }
      fn += 8.0 / 3.0 * r * r * r + 6.0 * r * r * nd;
    }
    bm += 8.0 * 3.0 * t.n * c;
    bs += we * 3.0 * t.n * c;
  }
  I.flops_leaf = fl; I.flops_reduce = fr; I.flops_node = fn; I.flops_bcast = fb;
  I.flops_factor = fl + fr + fn + fb;
  I.bytes_leaf = bl; I.bytes_reduce = br; I.bytes_bcast = bb;
  I.flops_matvec = fm; I.bytes_matvec = bm;
  I.flops_solve_only = fs; I.bytes_solve_only = bs;
  I.launches_factor = int(P->factor.l.size());
}

int check_system(const hodlr_system_t& s) {
  if (s.n <= 0 || s.threshold <= 0 || s.nrhs <= 0 || s.nranks <= 0 || !s.ranks) return HODLR_ERR_ARG;
  if (!s.X || !s.V || !s.K || !s.x) return HODLR_ERR_ARG;
  if (s.ldx < s.n || s.ldv < s.n || s.ldxs < s.n) return HODLR_ERR_ARG;
  if (s.klayout != HODLR_K_PACKED && s.klayout != HODLR_K_STACKED) return HODLR_ERR_ARG;
  if (s.klayout == HODLR_K_STACKED && s.ldk < s.n) return HODLR_ERR_ARG;
  if (s.n > (int64_t(1) << 31) - 1) return HODLR_ERR_UNSUPPORTED;
  return HODLR_OK;
}

void free_plan(hodlr_plan_s* P) {
  drop_graph(P);
  for (hipEvent_t e : P->pool) (void)hipEventDestroy(e);
  P->pool.clear();
  if (P->ws.p) (void)hipFree(P->ws.p);
  if (P->descs.p) (void)hipFree(P->descs.p);
  if (P->scratch.p) (void)hipFree(P->scratch.p);
}

int sync_stream(hodlr_plan_s* P) {
  return hipStreamSynchronize(P->stream) == hipSuccess ? HODLR_OK : HODLR_ERR_HIP;
}

int run_factor(hodlr_plan_s* P) {
  if (P->graphs()) {
    if (!P->g_factor.exec) {
      int rc = capture(P, P->g_factor, [&] { return run(P, P->factor); });
      if (rc) return rc;
    }
    if (replay(P, P->g_factor)) return HODLR_ERR_HIP;
  } else {
    int rc = run(P, P->factor);
    if (rc) return rc;
  }
  P->factored = true;
  return HODLR_OK;
}

// The refinement loop of hodlr_solve, driven from the device: the fused pass,
// the first residual and its check, then sweeps gated by RefineCtl::active,
// which the check after each residual clears once max ||r_j|| / ||b_j|| <=
// tol or maxiter sweeps are done.  With graphs on, one replay holds the pass
// and as many sweeps as the previous solve took, so a solve of the same kind
// of system is one graph launch and one host synchronisation; further sweeps
// (if still active) replay a one-sweep graph.
int run_solve(hodlr_plan_s* P, double tol, int maxiter) {
  RefineCtl* dctl = (RefineCtl*)(P->ws.p + P->ctl);
  P->ctl_host = RefineCtl{};
  P->ctl_host.tol = tol;
  P->ctl_host.maxiter = std::max(maxiter, 0);
  if (hipMemcpyAsync(dctl, &P->ctl_host, sizeof(RefineCtl), hipMemcpyHostToDevice, P->stream) != hipSuccess)
    return HODLR_ERR_HIP;
  int rc = HODLR_OK;
  if (P->graphs()) {
    const int S = std::min(P->sweep_hint, P->ctl_host.maxiter);
    if (!P->g_solve.exec || P->g_solve.sweeps != S) {
      rc = capture(P, P->g_solve, [&] {
        int r = run(P, P->factor);
        if (!r) r = run(P, P->first);
        for (int i = 0; i < S && !r; ++i) r = run(P, P->sweep);
        return r;
      });
      if (rc) return rc;
      P->g_solve.sweeps = S;
    }
    rc = replay(P, P->g_solve);
  } else {
    rc = run(P, P->factor);
    if (!rc) rc = run(P, P->first);
  }
  if (rc) return rc;
  P->factored = true;
  for (;;) {
    if (hipMemcpyAsync(&P->ctl_host, dctl, sizeof(RefineCtl), hipMemcpyDeviceToHost, P->stream) != hipSuccess ||
        sync_stream(P))
      return HODLR_ERR_HIP;
    if (!P->ctl_host.active) break;
    if (P->graphs()) {
      if (!P->g_sweep.exec && (rc = capture(P, P->g_sweep, [&] { return run(P, P->sweep); }))) return rc;
      rc = replay(P, P->g_sweep);
    } else {
      rc = run(P, P->sweep);
    }
    if (rc) return rc;
  }
  P->sweep_hint = P->ctl_host.iter;
  return HODLR_OK;
}

int read_relres(hodlr_plan_s* P, double* relres, double* mx) {
  double worst = 0;
  int cmax = 0;
  for (auto& S : P->sys) cmax = std::max(cmax, S.tree.nrhs);
  for (size_t s = 0; s < P->sys.size(); ++s) {
    SysState& S = P->sys[s];
    const int c = S.tree.nrhs;
    std::vector<double> rn(c), bn(c);
    if (hipMemcpyAsync(rn.data(), P->ws.p + S.rnorm, c * 8, hipMemcpyDeviceToHost, P->stream) != hipSuccess ||
        hipMemcpyAsync(bn.data(), P->ws.p + S.bnorm, c * 8, hipMemcpyDeviceToHost, P->stream) != hipSuccess)
      return HODLR_ERR_HIP;
    if (sync_stream(P)) return HODLR_ERR_HIP;
    for (int j = 0; j < c; ++j) {
      double v = bn[j] > 0 ? rn[j] / bn[j] : rn[j];
      if (!(v == v)) v = INFINITY;
      if (relres) relres[s * cmax + j] = v;
      worst = std::max(worst, v);
    }
  }
  if (mx) *mx = worst;
  return HODLR_OK;
}

// ---------------------------------------------------- fixed scratch buffer
// hodlr_dgesv_batched_noalloc keeps its descriptors and workspace in this
// file-scope device array instead of memory it allocates: one copy per GPU,
// held until the process ends.  Host code never uses the array's name as an
// address; it asks the runtime for this device's copy once and keeps it.
constexpr size_t kScratchBytes = size_t(64) << 20;      // 64 MiB per device
constexpr size_t kScratchDescBytes = size_t(4) << 20;   // of which the descriptor tables
constexpr int kScratchMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchMaxDevices] = {};

}  // namespace

static __device__ __align__(256) char scratch[kScratchBytes];

namespace {
// Device address of `scratch` on the current device (looked up on its first
// call on that device).
int scratch_on_current_device(char** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return HODLR_ERR_HIP;
  if (dev < 0 || dev >= kScratchMaxDevices) return HODLR_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return HODLR_ERR_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = (char*)g_scratch_ptr[dev];
  return HODLR_OK;
}
}  // namespace

// ======================================================================= C-ABI
extern "C" {

const char* hodlr_version(void) { return "fastsolver_amd 0.1 (gfx950)"; }

const char* hodlr_error_string(int32_t code) {
  switch (code) {
    case HODLR_OK: return "ok";
    case HODLR_ERR_ARG: return "invalid argument";
    case HODLR_ERR_TREE: return "invalid HODLR tree (sizes, ranks, or leaf rows <= rank)";
    case HODLR_ERR_UNSUPPORTED: return "leaf size or rank above the supported 1024";
    case HODLR_ERR_HIP: return "HIP runtime error";
    case HODLR_ERR_NOMEM: return "device allocation failed";
    case HODLR_ERR_SINGULAR: return "zero pivot in a leaf block or a Schur complement";
    case HODLR_ERR_NOTCONVERGED: return "refinement did not reach the tolerance";
    case HODLR_ERR_SCRATCH: return "input exceeds the fixed scratch buffer of the no-allocation entry point";
  }
  return "unknown error";
}

const char* hodlr_profile_category(int32_t i) { return (i >= 0 && i < CAT_N) ? kCatNames[i] : ""; }

int hodlr_plan_create(hodlr_plan_t* plan, const hodlr_system_t* systems, int32_t nsys, int32_t precision,
                      void* stream) {
  if (!plan || !systems || nsys <= 0) return HODLR_ERR_ARG;
  if (precision != HODLR_PREC_FP64 && precision != HODLR_PREC_FP32 && precision != HODLR_PREC_BF16)
    return HODLR_ERR_ARG;
  *plan = nullptr;
  std::unique_ptr<hodlr_plan_s> P(new hodlr_plan_s());
  P->precision = precision;
  P->wt = precision == HODLR_PREC_FP64 ? F64 : F32;
  P->mt = precision == HODLR_PREC_BF16 ? BF16 : P->wt;
  P->stream = (hipStream_t)stream;
  P->sys.resize(nsys);
  for (int s = 0; s < nsys; ++s) {
    int rc = check_system(systems[s]);
    if (rc) return rc;
    SysState& S = P->sys[s];
    S.s = systems[s];
    S.ranks.assign(systems[s].ranks, systems[s].ranks + systems[s].nranks);
    S.s.ranks = S.ranks.data();
    if (build_tree(S.s.n, S.s.threshold, S.s.nrhs, S.s.nranks, S.ranks.data(), S.tree) != TREE_OK)
      return HODLR_ERR_TREE;
    for (const TreeNode& x : S.tree.nodes) {
      if (x.is_leaf() && x.nrow > kMaxBlock) return HODLR_ERR_UNSUPPORTED;
      if (!x.is_leaf() && S.tree.rank(x.depth + 1) > kMaxBlock) return HODLR_ERR_UNSUPPORTED;
    }
    if (int64_t(S.tree.xcols()) * S.tree.n > (int64_t(1) << 40)) return HODLR_ERR_UNSUPPORTED;
    P->height = std::max(P->height, S.tree.height);
    P->nleaves += int(S.tree.leaves.size());
  }
  // workspace
  Layout L;
  if (layout_plan(P.get(), L)) return HODLR_ERR_ARG;
  size_t linfo = L.take(size_t(P->nleaves) * 4);
  size_t ninfo = L.take(size_t(std::max(P->nnodes, 1)) * 4);
  P->ctl = L.take(sizeof(RefineCtl));
  P->ws.bytes = L.top;
  if (hipMalloc((void**)&P->ws.p, P->ws.bytes) != hipSuccess) { free_plan(P.get()); return HODLR_ERR_NOMEM; }
  // a cleared workspace: padding rows that no kernel writes read as zero
  if (hipMemsetAsync(P->ws.p, 0, P->ws.bytes, P->stream) != hipSuccess ||
      hipStreamSynchronize(P->stream) != hipSuccess) {
    free_plan(P.get());
    return HODLR_ERR_HIP;
  }
  P->leaf_info = (int32_t*)(P->ws.p + linfo);
  P->node_info = (int32_t*)(P->ws.p + ninfo);
  // scratch must exist before descriptors reference it: size it with a dry run
  {
    Builder dry{P.get()};
    build_all(P.get(), dry);
    P->scratch_doubles = std::max<size_t>(dry.scratch_need, 1);
    if (hipMalloc((void**)&P->scratch.p, P->scratch_doubles * 8) != hipSuccess) {
      free_plan(P.get());
      return HODLR_ERR_NOMEM;
    }
    P->scratch.bytes = P->scratch_doubles * 8;
  }
  Builder B{P.get()};
  B.tunable = true;
  P->leaf = Proc(); P->extract = Proc(); P->residual = Proc(); P->update = Proc(); P->solve_only_leaf = Proc();
  P->matvec = Proc(); P->solve_out = Proc();
  build_all(P.get(), B);
  if (B.bad_gate) { free_plan(P.get()); return HODLR_ERR_UNSUPPORTED; }
  P->tile_choice.assign(B.ngid, -1);
  // the fused pass
  P->factor = Proc();
  P->factor.append(P->leaf);
  for (int h = 1; h <= P->height; ++h) {
    P->factor.append(P->reduce[h]);
    P->factor.append(P->node[h]);
    P->factor.append(P->bcast[h]);
  }
  P->factor.append(P->extract);
  P->descs.bytes = std::max<size_t>(B.blob.data.size(), 256);
  if (hipMalloc((void**)&P->descs.p, P->descs.bytes) != hipSuccess) { free_plan(P.get()); return HODLR_ERR_NOMEM; }
  // descriptors embed the final device addresses: upload once
  if (hipMemcpy(P->descs.p, B.blob.data.data(), B.blob.data.size(), hipMemcpyHostToDevice) != hipSuccess) {
    free_plan(P.get());
    return HODLR_ERR_HIP;
  }
  fill_info(P.get());
  *plan = P.release();
  return HODLR_OK;
}

int hodlr_plan_destroy(hodlr_plan_t P) {
  if (!P) return HODLR_ERR_ARG;
  (void)hipStreamSynchronize(P->stream);
  free_plan(P);
  delete P;
  return HODLR_OK;
}

int hodlr_plan_query(hodlr_plan_t P, hodlr_plan_info_t* info) {
  if (!P || !info) return HODLR_ERR_ARG;
  *info = P->info;
  return HODLR_OK;
}

int hodlr_factor_solve(hodlr_plan_t P) {
  if (!P) return HODLR_ERR_ARG;
  return run_factor(P);
}

int hodlr_leaf_solve(hodlr_plan_t P) { return P ? run(P, P->leaf) : HODLR_ERR_ARG; }
int hodlr_level_reduce(hodlr_plan_t P, int32_t h) {
  if (!P || h < 1 || h > P->height) return HODLR_ERR_ARG;
  return run(P, P->reduce[h]);
}
int hodlr_level_node_solve(hodlr_plan_t P, int32_t h) {
  if (!P || h < 1 || h > P->height) return HODLR_ERR_ARG;
  return run(P, P->node[h]);
}
int hodlr_level_broadcast(hodlr_plan_t P, int32_t h) {
  if (!P || h < 1 || h > P->height) return HODLR_ERR_ARG;
  int rc = run(P, P->bcast[h]);
  if (!rc && h == P->height) P->factored = true;
  return rc;
}
int hodlr_extract_solution(hodlr_plan_t P) { return P ? run(P, P->extract) : HODLR_ERR_ARG; }

static int solve_only_internal(hodlr_plan_t P) {
  int rc = run(P, P->solve_only_leaf);
  for (int h = 1; h <= P->height && !rc; ++h) {
    rc = run(P, P->reduce_s[h]);
    if (!rc) rc = run(P, P->node_s[h]);
    if (!rc) rc = run(P, P->bcast_s[h]);
  }
  return rc;
}

static int stage_copy(hodlr_plan_t P, const double* const* src, const int64_t* lds, bool to_rbuf,
                      double* const* dst, const int64_t* ldd) {
  // host-driven 2D copies between user arrays and the plan's residual buffer
  for (size_t s = 0; s < P->sys.size(); ++s) {
    SysState& S = P->sys[s];
    const size_t w = size_t(S.tree.n) * 8;
    hipError_t e;
    if (to_rbuf)
      e = hipMemcpy2DAsync(P->ws.p + S.rbuf, S.ldw * 8, src[s], lds[s] * 8, w, S.tree.nrhs, hipMemcpyDeviceToDevice,
                           P->stream);
    else
      e = hipMemcpy2DAsync(dst[s], ldd[s] * 8, P->ws.p + S.rbuf, S.ldw * 8, w, S.tree.nrhs, hipMemcpyDeviceToDevice,
                           P->stream);
    if (e != hipSuccess) return HODLR_ERR_HIP;
  }
  return HODLR_OK;
}

int hodlr_solve_only(hodlr_plan_t P, const double* const* rhs, const int64_t* ldr, double* const* out,
                     const int64_t* ldo) {
  if (!P || !rhs || !ldr || !out || !ldo) return HODLR_ERR_ARG;
  if (!P->factored) return HODLR_ERR_ARG;
  for (size_t s = 0; s < P->sys.size(); ++s)
    if (!rhs[s] || !out[s] || ldr[s] < P->sys[s].tree.n || ldo[s] < P->sys[s].tree.n) return HODLR_ERR_ARG;
  int rc = stage_copy(P, rhs, ldr, true, nullptr, nullptr);
  if (!rc) rc = solve_only_internal(P);
  if (!rc) rc = run(P, P->solve_out);              // rbuf <- yws (FP64)
  if (!rc) rc = stage_copy(P, nullptr, nullptr, false, out, ldo);
  return rc;
}

// ------------------------------------------------- standalone batched operators
// The BLAS/LAPACK calls inside the reference's task bodies, batched: every
// call builds its descriptors, runs on `stream` and synchronises it.
static int run_oneoff(hodlr_plan_s& P, Builder& B, const Proc& pr) {
  void* tmp = nullptr;
  if (hipMalloc(&tmp, std::max<size_t>(B.blob.data.size(), 256)) != hipSuccess) return HODLR_ERR_NOMEM;
  int rc = hipMemcpyAsync(tmp, B.blob.data.data(), B.blob.data.size(), hipMemcpyHostToDevice, P.stream) == hipSuccess
               ? HODLR_OK : HODLR_ERR_HIP;
  char* save = P.descs.p;
  P.descs.p = (char*)tmp;
  if (!rc) rc = run(&P, pr);
  P.descs.p = save;
  if (hipStreamSynchronize(P.stream) != hipSuccess && !rc) rc = HODLR_ERR_HIP;
  (void)hipFree(tmp);
  return rc;
}

static int gemm_batched(DT t, DT mt, int32_t batch, const char* transa, const int32_t* m, const int32_t* n, const int32_t* k,
                        const double* alpha, const void* const* A, const int64_t* lda, const void* const* B,
                        const int64_t* ldb, const double* beta, void* const* C, const int64_t* ldc, void* stream) {
  if (batch <= 0 || !transa || !m || !n || !k || !alpha || !A || !lda || !B || !ldb || !beta || !C || !ldc)
    return HODLR_ERR_ARG;
  hodlr_plan_s P;
  P.stream = (hipStream_t)stream;
  Builder Bd{&P};
  Proc pr;
  const size_t es = esize(t);
  for (int tr = 0; tr < 2; ++tr) {
    std::vector<Builder::G> gs;
    for (int i = 0; i < batch; ++i) {
      const bool T = transa[i] == 't' || transa[i] == 'T';
      if (!T && transa[i] != 'n' && transa[i] != 'N') return HODLR_ERR_ARG;
      if (T != bool(tr) || m[i] <= 0 || n[i] <= 0) continue;
      if (k[i] < 0 || lda[i] < 1 || ldb[i] < std::max(k[i], 1) || ldc[i] < m[i]) return HODLR_ERR_ARG;
      if (lda[i] < (T ? std::max(k[i], 1) : m[i])) return HODLR_ERR_ARG;
      Builder::G g{};
      g.A = A[i]; g.lda = lda[i]; g.B = B[i]; g.ldb = ldb[i];
      g.Cin = C[i]; g.ldcin = ldc[i]; g.Cout = C[i]; g.ldcout = ldc[i];
      g.m = m[i]; g.n = n[i]; g.k = k[i]; g.alpha = alpha[i]; g.beta = beta[i];
      gs.push_back(g);
    }
    if (gs.empty()) continue;
    if (t == F32 && tr) return HODLR_ERR_UNSUPPORTED;   // fp32 products are NN only
    Bd.gemm(pr, CAT_OTHER, t, t, t, mt, tr == 1, tr == 1 ? 1 : (t == F32 ? 3 : 0), gs);
  }
  (void)es;
  return run_oneoff(P, Bd, pr);
}

int hodlr_dgemm_batched(int32_t batch, const char* transa, const int32_t* m, const int32_t* n, const int32_t* k,
                        const double* alpha, const double* const* A, const int64_t* lda, const double* const* B,
                        const int64_t* ldb, const double* beta, double* const* C, const int64_t* ldc, void* stream) {
  return gemm_batched(F64, F64, batch, transa, m, n, k, alpha, (const void* const*)A, lda, (const void* const*)B, ldb, beta,
                      (void* const*)C, ldc, stream);
}

int hodlr_sgemm_batched(int32_t batch, const int32_t* m, const int32_t* n, const int32_t* k, const double* alpha,
                        const float* const* A, const int64_t* lda, const float* const* B, const int64_t* ldb,
                        const double* beta, float* const* C, const int64_t* ldc, void* stream) {
  if (batch <= 0) return HODLR_ERR_ARG;
  std::vector<char> tn(batch, 'n');
  return gemm_batched(F32, F32, batch, tn.data(), m, n, k, alpha, (const void* const*)A, lda, (const void* const*)B,
                      ldb, beta, (void* const*)C, ldc, stream);
}

int hodlr_sgemm_batched_bf16(int32_t batch, const int32_t* m, const int32_t* n, const int32_t* k,
                             const double* alpha, const float* const* A, const int64_t* lda,
                             const float* const* B, const int64_t* ldb, const double* beta, float* const* C,
                             const int64_t* ldc, void* stream) {
  if (batch <= 0) return HODLR_ERR_ARG;
  std::vector<char> tn(batch, 'n');
  return gemm_batched(F32, BF16, batch, tn.data(), m, n, k, alpha, (const void* const*)A, lda,
                      (const void* const*)B, ldb, beta, (void* const*)C, ldc, stream);
}

// hodlr_dgesv_batched's launches on a workspace at `ws` (dgesv_layout's
// offsets): LU with partial pivoting of every A_i in place, B_i copied aside,
// then B_i = U^-1 L^-1 P (copy).
struct DgesvLayout {
  std::vector<size_t> piv, prm, dv, bc;
  size_t inf = 0, bytes = 0;
};

static int dgesv_check(int32_t batch, const int32_t* n, const int32_t* nrhs, double* const* A, const int64_t* lda,
                       double* const* B, const int64_t* ldb) {
  if (batch <= 0 || !n || !nrhs || !A || !lda || !B || !ldb) return HODLR_ERR_ARG;
  for (int i = 0; i < batch; ++i) {
    if (n[i] <= 0 || nrhs[i] < 0 || lda[i] < n[i] || ldb[i] < n[i] || !A[i] || !B[i]) return HODLR_ERR_ARG;
    if (n[i] > kMaxBlock) return HODLR_ERR_UNSUPPORTED;
  }
  return HODLR_OK;
}

static DgesvLayout dgesv_layout(int32_t batch, const int32_t* n, const int32_t* nrhs) {
  // pivots, permutations, inverted diagonal blocks, a copy of B, info
  DgesvLayout D;
  Layout L;
  D.piv.resize(batch); D.prm.resize(batch); D.dv.resize(batch); D.bc.resize(batch);
  for (int i = 0; i < batch; ++i) {
    D.piv[i] = L.take(size_t(n[i]) * 4);
    D.prm[i] = L.take(size_t(n[i]) * 4);
    D.dv[i] = L.take(size_t(cdiv(n[i], 64)) * 2 * 64 * 64 * 8);
    D.bc[i] = L.take(size_t(n[i]) * std::max(nrhs[i], 1) * 8);
  }
  D.inf = L.take(size_t(batch) * 4);
  D.bytes = L.top;
  return D;
}

static void dgesv_build(hodlr_plan_s& P, Builder& Bd, Proc& pr, const DgesvLayout& D, int32_t batch,
                        const int32_t* n, const int32_t* nrhs, double* const* A, const int64_t* lda,
                        double* const* B, const int64_t* ldb) {
  std::vector<LuDesc> lus;
  std::vector<CopyDesc> cps;
  std::vector<TrsmDesc> trs;
  for (int i = 0; i < batch; ++i) {
    LuDesc l{};
    l.A = A[i]; l.lda = lda[i]; l.m = n[i];
    l.ipiv = (int32_t*)(P.ws.p + D.piv[i]);
    l.perm = (int32_t*)(P.ws.p + D.prm[i]);
    l.info = (int32_t*)(P.ws.p + D.inf) + i;
    l.dinv = P.ws.p + D.dv[i];
    lus.push_back(l);
    CopyDesc c{};
    c.src = B[i]; c.lds = ldb[i]; c.dst = P.ws.p + D.bc[i]; c.ldd = n[i];
    c.rows = n[i]; c.cols = nrhs[i]; c.alpha = 1.0;
    cps.push_back(c);
    TrsmDesc t{};
    t.LU = A[i]; t.ldlu = lda[i]; t.dinv = l.dinv; t.perm = l.perm;
    t.Xin = P.ws.p + D.bc[i]; t.ldin = n[i]; t.Xout = B[i]; t.ldout = ldb[i];
    t.m = n[i]; t.ncols = nrhs[i];
    trs.push_back(t);
  }
  Bd.lu(pr, CAT_OTHER, CAT_OTHER, CAT_OTHER, CAT_OTHER, F64, lus);
  Bd.copy(pr, CAT_OTHER, F64, F64, cps);
  Bd.trsm(pr, CAT_OTHER, F64, F64, F64, trs);
}

int hodlr_dgesv_batched(int32_t batch, const int32_t* n, const int32_t* nrhs, double* const* A, const int64_t* lda,
                        double* const* B, const int64_t* ldb, int32_t* info, void* stream) {
  int rc = dgesv_check(batch, n, nrhs, A, lda, B, ldb);
  if (rc) return rc;
  hodlr_plan_s P;
  P.stream = (hipStream_t)stream;
  const DgesvLayout D = dgesv_layout(batch, n, nrhs);
  if (hipMalloc((void**)&P.ws.p, D.bytes) != hipSuccess) return HODLR_ERR_NOMEM;
  P.ws.bytes = D.bytes;
  Builder Bd{&P};
  Proc pr;
  dgesv_build(P, Bd, pr, D, batch, n, nrhs, A, lda, B, ldb);
  rc = run_oneoff(P, Bd, pr);
  if (!rc && info && hipMemcpy(info, P.ws.p + D.inf, size_t(batch) * 4, hipMemcpyDeviceToHost) != hipSuccess)
    rc = HODLR_ERR_HIP;
  (void)hipFree(P.ws.p);
  P.ws.p = nullptr;
  return rc;
}

int hodlr_dgesv_batched_noalloc(int32_t batch, const int32_t* n, const int32_t* nrhs, double* const* A,
                                const int64_t* lda, double* const* B, const int64_t* ldb, int32_t* info,
                                void* stream) {
  int rc = dgesv_check(batch, n, nrhs, A, lda, B, ldb);
  if (rc) return rc;
  const DgesvLayout D = dgesv_layout(batch, n, nrhs);
  if (D.bytes > kScratchBytes - kScratchDescBytes) return HODLR_ERR_SCRATCH;
  char* base = nullptr;
  if ((rc = scratch_on_current_device(&base))) return rc;
  hodlr_plan_s P;
  P.stream = (hipStream_t)stream;
  P.ws.p = base + kScratchDescBytes;   // workspace after the descriptor region
  P.ws.bytes = D.bytes;
  Builder Bd{&P};
  Proc pr;
  dgesv_build(P, Bd, pr, D, batch, n, nrhs, A, lda, B, ldb);
  if (Bd.blob.data.size() > kScratchDescBytes) return HODLR_ERR_SCRATCH;
  P.descs.p = base;
  if (hipMemcpyAsync(base, Bd.blob.data.data(), Bd.blob.data.size(), hipMemcpyHostToDevice, P.stream) != hipSuccess)
    return HODLR_ERR_HIP;
  rc = run(&P, pr);
  if (!rc && info &&
      hipMemcpyAsync(info, P.ws.p + D.inf, size_t(batch) * 4, hipMemcpyDeviceToHost, P.stream) != hipSuccess)
    rc = HODLR_ERR_HIP;
  // the descriptor blob must stay alive until its upload is done, and info is
  // valid on return
  if (hipStreamSynchronize(P.stream) != hipSuccess && !rc) rc = HODLR_ERR_HIP;
  P.ws.p = nullptr;
  P.descs.p = nullptr;
  return rc;
}

int hodlr_matvec(hodlr_plan_t P, const double* const* x, const int64_t* ldx, double* const* y,
                 const int64_t* ldy) {
  if (!P || !x || !ldx || !y || !ldy) return HODLR_ERR_ARG;
  for (size_t s = 0; s < P->sys.size(); ++s) {
    SysState& S = P->sys[s];
    if (!x[s] || !y[s] || ldx[s] < S.tree.n || ldy[s] < S.tree.n) return HODLR_ERR_ARG;
    const size_t w = size_t(S.tree.n) * 8;
    if (hipMemcpy2DAsync(P->ws.p + S.mvx, S.ldw * 8, x[s], ldx[s] * 8, w, S.tree.nrhs, hipMemcpyDeviceToDevice,
                         P->stream) != hipSuccess)
      return HODLR_ERR_HIP;
  }
  int rc = run(P, P->matvec);                       // rbuf <- A mvx
  if (!rc) rc = stage_copy(P, nullptr, nullptr, false, y, ldy);
  return rc;
}

int hodlr_residual(hodlr_plan_t P, double* relres, double* relres_max) {
  if (!P) return HODLR_ERR_ARG;
  int rc = run(P, P->residual);
  if (rc) return rc;
  return read_relres(P, relres, relres_max);
}

int hodlr_solve(hodlr_plan_t P, double tol, int32_t maxiter, hodlr_solve_report_t* rep) {
  if (!P) return HODLR_ERR_ARG;
  hodlr_solve_report_t R{};
  int rc = run_solve(P, tol, maxiter);
  if (rc) return rc;
  R.relres_first = P->ctl_host.relres_first;
  R.iterations = P->ctl_host.iter;
  R.relres = P->ctl_host.relres;
  R.converged = R.relres <= tol;
  if (rep) *rep = R;
  return R.converged ? HODLR_OK : HODLR_ERR_NOTCONVERGED;
}

int hodlr_status(hodlr_plan_t P, int32_t* leaf_info, int32_t* node_info, int32_t* n_nodes) {
  if (!P) return HODLR_ERR_ARG;
  if (sync_stream(P)) return HODLR_ERR_HIP;
  std::vector<int32_t> li(P->nleaves), ni(std::max(P->nnodes, 1));
  if (hipMemcpy(li.data(), P->leaf_info, li.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return HODLR_ERR_HIP;
  if (P->nnodes && hipMemcpy(ni.data(), P->node_info, size_t(P->nnodes) * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return HODLR_ERR_HIP;
  if (leaf_info) memcpy(leaf_info, li.data(), li.size() * 4);
  if (node_info && P->nnodes) memcpy(node_info, ni.data(), size_t(P->nnodes) * 4);
  if (n_nodes) *n_nodes = P->nnodes;
  for (int v : li) if (v) return HODLR_ERR_SINGULAR;
  for (int i = 0; i < P->nnodes; ++i) if (ni[i]) return HODLR_ERR_SINGULAR;
  return HODLR_OK;
}

int hodlr_workspace(hodlr_plan_t P, int32_t sys, void** ptr, int64_t* ld, int32_t* dtype) {
  if (!P || sys < 0 || sys >= int(P->sys.size())) return HODLR_ERR_ARG;
  if (ptr) *ptr = P->ws.p + P->sys[sys].xws;
  if (ld) *ld = P->sys[sys].ldw;
  if (dtype) *dtype = P->wt == F64 ? 0 : 1;
  return HODLR_OK;
}

int hodlr_use_graph(hodlr_plan_t P, int32_t enable) {
  if (!P) return HODLR_ERR_ARG;
  if (enable && P->stream == nullptr) return HODLR_ERR_ARG;
  P->use_graph = enable != 0;
  return HODLR_OK;
}

int hodlr_autotune(hodlr_plan_t P, int32_t reps, int32_t* switched) {
  if (!P) return HODLR_ERR_ARG;
  const size_t ng = P->tile_choice.size();
  if (switched) *switched = 0;
  if (!ng) return HODLR_OK;
  for (int c = 0; c < 2; ++c) P->tune_ms[c].assign(ng, 0.0);
  int rc = HODLR_OK;
  const int R = std::max(reps, 1) + 1;   // the first round only warms up
  for (int r = 0; r < R && !rc; ++r) {
    for (int cfg = 0; cfg < 2 && !rc; ++cfg) {
      P->tuning = true;
      P->tune_cfg = cfg;
      rc = run(P, P->factor);
      if (!rc) rc = solve_only_internal(P);
      P->tuning = false;
      if (hipStreamSynchronize(P->stream) != hipSuccess && !rc) rc = HODLR_ERR_HIP;
      for (auto& q : P->tune_pending) {
        float ms = 0;
        if (!rc && hipEventElapsedTime(&ms, q.a, q.b) != hipSuccess) rc = HODLR_ERR_HIP;
        if (r > 0) P->tune_ms[cfg][q.gid] += ms;
        P->pool.push_back(q.a);
        P->pool.push_back(q.b);
      }
      P->tune_pending.clear();
    }
  }
  if (rc) return rc;
  int sw = 0;
  for (size_t g = 0; g < ng; ++g) {
    int best = P->tune_ms[1][g] < P->tune_ms[0][g] ? 1 : 0;
    P->tile_choice[g] = best;
    sw += best;
  }
  if (switched) *switched = sw;
  drop_graph(P);
  P->factored = false;   // the tuning runs left the workspace in an arbitrary state
  return HODLR_OK;
}

int hodlr_tune_force(hodlr_plan_t P, int32_t cfg) {
  if (!P || cfg < -1 || cfg > 1) return HODLR_ERR_ARG;
  for (auto& c : P->tile_choice) c = cfg;
  drop_graph(P);
  return HODLR_OK;
}

int hodlr_profile_enable(hodlr_plan_t P, int32_t mask) {
  if (!P) return HODLR_ERR_ARG;
  // (captured graphs stay; they are bypassed while the mask is non-zero)
  P->prof_mask = uint32_t(mask);
  return HODLR_OK;
}

int hodlr_profile_reset(hodlr_plan_t P) {
  if (!P) return HODLR_ERR_ARG;
  int rc = harvest(P);
  for (int i = 0; i < CAT_N; ++i) { P->prof_ms[i] = 0; P->prof_n[i] = 0; P->prof_flops[i] = 0; P->prof_bytes[i] = 0; }
  return rc;
}

int hodlr_profile_read(hodlr_plan_t P, double* times_ms, int32_t* launches, double* flops, double* bytes) {
  if (!P) return HODLR_ERR_ARG;
  int rc = harvest(P);
  if (rc) return rc;
  for (int i = 0; i < CAT_N; ++i) {
    if (times_ms) times_ms[i] = P->prof_ms[i];
    if (launches) launches[i] = P->prof_n[i];
    if (flops) flops[i] = P->prof_flops[i];
    if (bytes) bytes[i] = P->prof_bytes[i];
  }
  return HODLR_OK;
}

}  // extern "C"

// The reference's single_launch driver (test/single_launch.cc:16-77) on one
// GPU through the C++ facade: 3 levels, rank 30, threshold 60 (N = 480,
// 8 leaves), 2 right-hand sides, diagonal 1e4, seed 1245667.
//
//   built by fastsolver_amd/build.py: g++ -std=c++17 -I include single_launch.cc
//   -L fastsolver_amd -lfastsolver_amd -lfastsolver_amd_gen -Wl,-rpath,$ORIGIN/..
#include <cstdio>
#include <cstdlib>

#include "fastsolver_amd.hpp"

using namespace fastsolver_amd;

static int run(int argc, char** argv) {
  int gloTreeLevel = 3, numMachineNodes = 2;
  long seed = 1245667;
  int nRHS = 2, rank = 30, threshold = 60, leafSize = 1;
  double diagonal = 1.0e4;
  int precision = argc > 1 ? atoi(argv[1]) : HODLR_PREC_FP64;
  const bool rhs_last = argc > 2 && atoi(argv[2]) != 0;   // init_rhs after init_circulant_matrix

  register_solver_tasks();
  int gloLevel = gloTreeLevel, subLevel = gloLevel;
  int nRow = threshold * (1 << subLevel);
  Range procs(numMachineNodes);
  HodlrMatrix hMatrix(nRHS, nRow, gloLevel, subLevel, rank, threshold, leafSize, "global");
  hMatrix.create_tree();
  if (!rhs_last) hMatrix.init_rhs(seed, procs);
  hMatrix.init_circulant_matrix(diagonal, procs);
  if (rhs_last) hMatrix.init_rhs(seed, procs);

  FastSolver fs(precision);
  fs.bfs_solve(hMatrix, procs);

  // the tree the solve ran on (hodlr_matrix.h:49-50)
  int depth = 0;
  const Node* leaf = hMatrix.uroot;
  for (; !leaf->is_legion_leaf(); leaf = leaf->lchild) ++depth;
  std::cout << "uroot: " << hMatrix.uroot->nrow << " x " << hMatrix.uroot->ncol << ", " << depth
            << " levels; leftmost leaf " << leaf->nrow << " rows, U block at column " << leaf->col_beg
            << "; vroot depth-1 block at V column " << hMatrix.vroot->lchild->col_beg << std::endl;

  std::cout << "\n================================" << std::endl;
  std::cout << "problem information:" << std::endl;
  int nleaf = hMatrix.get_num_leaf();
  std::cout << "  Legion leaf : " << nleaf << std::endl
            << "  Legion leaf / node: " << nleaf / procs.size() << std::endl;
  hMatrix.display_launch_time();
  fs.display_launch_time();
  double rr = fs.residual();
  std::printf("Residual: %.3e\n", rr);
  int nregion = nleaf;
  double err = compute_L2_error(hMatrix, seed, nRow, nregion, nRHS, rank, diagonal);
  std::cout << "================================\n" << std::endl;
  return (rr <= 1e-12 && err <= 1e-12) ? 0 : 1;
}

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "single_launch: %s\n", e.what());
    return 2;
  }
}

// HODLR cluster tree and column layout, shared by the device library's plan
// builder and the host input generator.
//
// Split rule (src/htree/hodlr_matrix.cc:138-169): a node with nrow > threshold
// gets a left child of nrow/2 rows and a right child of nrow - nrow/2 rows.
// Column layout (hodlr_matrix.cc:48,151-155; node.cc:46-57): a node at depth
// d >= 1 owns ncol = r_d columns starting at col_beg = c + r_1 + ... + r_{d-1}
// of the right-hand-side array X = [b | U(1) | U(2) | ...].  The V array holds
// the same blocks without the c leading columns.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <vector>

namespace hodlr {

struct TreeNode {
  int64_t row_beg = 0;
  int64_t nrow = 0;
  int depth = 0;
  int parent = -1;
  int child[2] = {-1, -1};
  int height = 0;   // 0 for leaves, 1 + max(child heights) otherwise
  int leaf_id = -1; // index in row order, leaves only
  bool is_leaf() const { return child[0] < 0; }
};

struct Tree {
  int64_t n = 0;
  int64_t threshold = 0;
  int nrhs = 0;
  int maxdepth = 0;
  int height = 0;
  std::vector<TreeNode> nodes;   // breadth-first; nodes[0] is the root
  std::vector<int> leaves;       // node ids in row order
  std::vector<int64_t> rsum;     // rsum[d] = r_1 + ... + r_d, rsum[0] = 0

  int rank(int d) const { return int(rsum[d] - rsum[d - 1]); }
  int64_t col_beg(int d) const { return nrhs + rsum[d - 1]; }        // depth d >= 1
  int64_t xcols() const { return nrhs + rsum[maxdepth]; }           // columns of X
  int64_t vcols() const { return rsum[maxdepth]; }                  // columns of V
  int64_t leaf_nrhs(int id) const {                                 // dgesv NRHS of a leaf
    const TreeNode& t = nodes[id];
    return t.depth == 0 ? nrhs : col_beg(t.depth) + rank(t.depth);
  }
};

// Error codes shared with the C-ABI.
enum TreeError { TREE_OK = 0, TREE_BAD_SIZE = -101, TREE_BAD_RANK = -102, TREE_LEAF_TOO_SMALL = -103 };

// ranks: either one entry (same rank at every depth, as the reference) or one
// entry per depth 1..maxdepth.
inline int build_tree(int64_t n, int64_t threshold, int nrhs, int nranks, const int* ranks, Tree& t) {
  t = Tree();
  if (n <= 0 || threshold <= 0 || nrhs <= 0 || nranks <= 0) return TREE_BAD_SIZE;
  t.n = n;
  t.threshold = threshold;
  t.nrhs = nrhs;
  TreeNode root;
  root.nrow = n;
  t.nodes.push_back(root);
  for (size_t i = 0; i < t.nodes.size(); ++i) {
    if (t.nodes[i].nrow <= threshold) continue;
    TreeNode a, b;
    a.row_beg = t.nodes[i].row_beg;
    a.nrow = t.nodes[i].nrow / 2;
    b.row_beg = a.row_beg + a.nrow;
    b.nrow = t.nodes[i].nrow - a.nrow;
    a.depth = b.depth = t.nodes[i].depth + 1;
    a.parent = b.parent = int(i);
    t.nodes[i].child[0] = int(t.nodes.size());
    t.nodes.push_back(a);
    t.nodes[i].child[1] = int(t.nodes.size());
    t.nodes.push_back(b);
  }
  for (const TreeNode& x : t.nodes) t.maxdepth = std::max(t.maxdepth, x.depth);
  if (nranks != 1 && nranks < t.maxdepth) return TREE_BAD_RANK;
  t.rsum.assign(t.maxdepth + 1, 0);
  for (int d = 1; d <= t.maxdepth; ++d) {
    int r = nranks == 1 ? ranks[0] : ranks[d - 1];
    if (r <= 0) return TREE_BAD_RANK;
    t.rsum[d] = t.rsum[d - 1] + r;
  }
  for (int i = int(t.nodes.size()) - 1; i >= 0; --i) {
    TreeNode& x = t.nodes[i];
    if (!x.is_leaf())
      x.height = 1 + std::max(t.nodes[x.child[0]].height, t.nodes[x.child[1]].height);
  }
  t.height = t.nodes[0].height;
  std::vector<int> stack{0};
  while (!stack.empty()) {
    int i = stack.back();
    stack.pop_back();
    TreeNode& x = t.nodes[i];
    if (x.is_leaf()) {
      x.leaf_id = int(t.leaves.size());
      t.leaves.push_back(i);
      // the reference asserts leaf rows > rank (hodlr_matrix.cc:167)
      if (x.depth >= 1 && x.nrow <= t.rank(x.depth)) return TREE_LEAF_TOO_SMALL;
    } else {
      stack.push_back(x.child[1]);
      stack.push_back(x.child[0]);
    }
  }
  return TREE_OK;
}

}  // namespace hodlr

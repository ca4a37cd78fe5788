// Octree and U/V/W/X interaction lists for one rank (the single-rank part of the
// reference's Let3d_MPI, let3d_mpi.hpp/.cpp), built on the host at setup and
// flattened into CSR arrays that are uploaded to HBM once.
//
// Node numbering, child order, tags and list contents reproduce the reference:
//   * leaves: a node splits iff it holds more than ptsMax distinct depth-maxLevel
//     cells and is above maxLevel (Dendro ot::points2Octree, let3d_mpi.cpp:268);
//     the root always splits (let3d_mpi.cpp:283-288);
//   * numbering: CreateAncestors (let3d_mpi.cpp:611-726) -- root 0, child blocks
//     appended in Dendro Morton order (z most significant); child = 4x+2y+z;
//   * point -> leaf by x >= centre tests (let3d_mpi.cpp:1255-1267);
//   * tags and packed ranges (srcData :1209-1465, trgData :1470-1631);
//   * lists (calGlbNodeLists :1637-1768), ascending node index.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace kifmm {

enum NodeTag { TAG_CBTR = 1, TAG_USER = 2, TAG_EVTR = 4, TAG_OWNR = 8, TAG_SRCE = 16 };
enum ListKind { LIST_U = 0, LIST_V = 1, LIST_W = 2, LIST_X = 3 };

struct Tree {
  int n_nodes = 0, max_level = 0, root_level = 1, max_depth = 0;
  double ctr[3] = {0.5, 0.5, 0.5};
  double rad0 = 0.5;
  std::vector<int> parent, child, depth, path, tag;
  std::vector<int> src_cnt, trg_cnt;    // points in each subtree
  std::vector<int> src_beg, src_num;    // packed source range of owned leaves, else -1
  std::vector<int> trg_beg, trg_num;    // packed target range of evaluator leaves, else -1
  std::vector<int> src_perm, trg_perm;  // packed slot -> caller's point index
  std::vector<int64_t> off[4];          // CSR offsets per list kind (n_nodes + 1)
  std::vector<int> lst[4];

  bool leaf(int i) const { return child[i] < 0; }
  double radius(int i) const;           // let3d_mpi.cpp:1824-1827
  void center(int i, double* c) const;  // let3d_mpi.cpp:1808-1818
};

// Returns 0 on success; on failure sets *err and returns a negative code.
int build_tree(const double* spos, int64_t ns, const double* tpos, int64_t nt, const double ctr[3],
               int root_level, int pts_max, int max_level, Tree& t, std::string* err);

// Forest of independent problems: the trees of parts[0..P) in one numbering.
// Problem p's nodes follow those of problems 0..p-1 (node offset = their node
// count), and its sources and targets follow theirs in the packed and in the
// caller's order (src_off[p], trg_off[p]: the points of problems before p).
// Lists, tags and ranges keep their meaning inside each tree; no list crosses
// a tree.  The parts share ctr, root level and maxLevel.  parts is consumed.
void merge_trees(std::vector<Tree>& parts, const std::vector<int64_t>& src_off, const std::vector<int64_t>& trg_off,
                 Tree& out);

}  // namespace kifmm

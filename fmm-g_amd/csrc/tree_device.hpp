// The octree and lists of let3d.hpp built on the GPU (tree_device.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "let3d.hpp"

namespace kifmm {

// The Tree that build_tree gives for one problem (nprob == 1), or that
// merge_trees gives over build_tree of each problem (a forest), computed on the
// device and copied into `t`.  ns[p] / nt[p]: sources and targets of problem p;
// spos / tpos (host) hold the problems one after another.  Same argument checks
// and error texts as the host builder.  ms (optional, 2 entries): build time on
// the device and copy-back time, milliseconds.
int build_tree_device(int nprob, const int64_t* ns, const double* spos, const int64_t* nt, const double* tpos,
                      const double ctr[3], int root_level, int pts_max, int max_level, Tree& t, std::string* err,
                      hipStream_t s, double* ms);

}  // namespace kifmm

#pragma once
#include "fmm_common.hpp"

namespace fmm {

struct TreeScratch {
    DevBuf flag, pos, mask, key, key2, val, sflag, srank, total, tmp;
    std::vector<DevBuf> lvl_idx, lvl_node;   // per level: point index and node of every arrival
    std::vector<int> lvl_n;
};

// Builds the reference octree of the n points d_xyz (double3, device) rooted at the box
// (root_c, root_h).  thr: a node subdivides when its arrivals reach thr.  cubic_scale:
// expansion scale per level = cube half size (else |half|/sqrt(3)).
void build_tree(TreeDev& T, TreeScratch& S, const double* d_xyz, int n, const double root_c[3],
                const double root_h[3], int thr, int max_depth, bool cubic_scale, hipStream_t st);

}  // namespace fmm

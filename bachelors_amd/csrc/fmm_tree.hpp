#pragma once
#include "fmm_common.hpp"

namespace fmm {

struct TreeScratch {
    DevBuf flag, pos, mask, key, key2, val, sflag, srank, total, tmp;
    DevBuf k64, k64c, k64s, irr, order, node_of, sb, se, need, dlevel, partial;   // one-sort build
    uint32_t h_last[4] = {0, 0, 0, 0};
    std::vector<DevBuf> lvl_idx, lvl_node;   // per level: point index and node of every arrival
    std::vector<int> lvl_n;
};

// Builds the reference octree of the n points d_xyz (double3, device) rooted at the box
// (root_c, root_h).  thr: a node subdivides when its arrivals reach thr.  cubic_scale:
// expansion scale per level = the largest half size (else |half|/sqrt(3)).  One radix sort by
// point path when every point's child is unique down to its leaf; otherwise (gaps or
// overlaps of the float-rounded child boxes) level by level.  T.sorted_build tells which.
void build_tree(TreeDev& T, TreeScratch& S, const double* d_xyz, int n, const double root_c[3],
                const double root_h[3], int thr, int max_depth, bool cubic_scale, hipStream_t st);

// copies the node ranges, levels, parents and centres to the host mirrors (once per build;
// synchronises the device)
void fetch_tree_details(TreeDev& T);

}  // namespace fmm

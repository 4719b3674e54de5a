#pragma once
#include <functional>

#include "fmm_common.hpp"

namespace fmm {

struct ListScratch {
    DevBuf far, close, frontA, frontB, counters, key, val, key2, val2, flag, pos, tmp, hs, srec, trec;
    DevBuf perm, perm2, wkey, wkey2, tkey, tkey2, idx;   // wide class codes (deep adaptive trees)
    unsigned far_cap = 0, close_cap = 0, front_cap = 0;
    unsigned h_seed[3] = {0, 0, 0};        // host staging of the root pair and its count
    double h_half[6 * MAX_LEVELS] = {0};   // host staging of the per-level half sizes [tree][level][axis]
};

// Far pairs sorted by (class, target) -- class = (level difference, centre offset in
// units of the finer half size) in the shared frame -- and close pairs sorted by
// (target, source) with per-target ranges.
struct Lists {
    int num_far = 0, num_close = 0, num_classes = 0, bfs_steps = 0;
    DevBuf far_t, far_s, far_code;          // int, int, uint32 class code
    DevBuf class_begin;                     // int[num_classes]
    std::vector<int> h_class_begin;         // num_classes + 1
    DevBuf close_t, close_s;                // int
    DevBuf close_begin, close_end;          // per target node
};

// far_ready (may be null) is recorded on st once the far pairs and classes are complete, then
// after_far (may be empty) is called (the host-side far data, num_far, num_classes and
// h_class_begin, is final), then the close pairs are sorted.  Returns once everything is queued
// on st (the host waits only for the list counts it reads back).
void build_lists(Lists& out, ListScratch& L, const TreeDev& S, const TreeDev& T, double mac_scale, bool classes,
                 hipStream_t st, hipEvent_t far_ready = nullptr, const std::function<void()>& after_far = {});

}  // namespace fmm

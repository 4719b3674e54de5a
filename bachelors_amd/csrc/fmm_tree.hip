// Octree build on the GPU, level by level, with the reference's exact rules:
//  - half-open boxes [c-h, c+h) compared in double (box.cpp:13-24);
//  - a node subdivides when its arrivals reach `thr` (source tree thr = cap,
//    quadrature_octree.cpp:38; target tree thr = cap + 1, calculation_point_octree.cpp:41);
//  - 8 children in reference order with centres rounded through float
//    (quadrature_octree.cpp:74-97) -- the child boxes of a non-dyadic root can leave
//    one-ulp gaps (points dropped) or overlaps (points inserted twice); both are kept;
//  - inside a leaf the points keep insertion (input) order.
// A point's path is computed from the actual float-rounded child boxes, so the tree is
// identical to the one the reference's pointer insertion builds, in canonical BFS node
// numbering, with the leaf contents laid out in DFS order ("slots").
#include <cmath>

#include "fmm_prims.hpp"
#include "fmm_tree.hpp"

namespace fmm {

static const int CHILD_SIGN[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                     {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
__constant__ int c_child_sign[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                       {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

__device__ __forceinline__ bool in_range(double v, double c, double h) {
    return __dsub_rn(c, h) <= v && v < __dadd_rn(c, h);
}

__global__ void k_root_filter(const double* __restrict__ P, int n, double3 c, double3 h, uint32_t* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = P[3 * i], y = P[3 * i + 1], z = P[3 * i + 2];
    flag[i] = (in_range(x, c.x, h.x) && in_range(y, c.y, h.y) && in_range(z, c.z, h.z)) ? 1u : 0u;
}

__global__ void k_compact_root(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, int n,
                               int* __restrict__ idx, int* __restrict__ node) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    idx[pos[i]] = i;
    node[pos[i]] = 0;
}

// subdivide flags of the level's nodes
__global__ void k_subdiv_flags(const int* __restrict__ arrivals, int lb, int nnodes, int thr, int allow,
                               uint32_t* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    flag[i] = (allow && arrivals[lb + i] >= thr) ? 1u : 0u;
}

// create the 8 children of every subdivided node (centres through float, quadrature_octree.cpp:76-96)
__global__ void k_make_children(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ rank, int lb,
                                int nnodes, int child_lb, int child_level, double3 ph,
                                double* __restrict__ center, int* __restrict__ first_child,
                                int* __restrict__ parent, int* __restrict__ level, int* __restrict__ arrivals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    int node = lb + i;
    if (!flag[i]) { first_child[node] = -1; return; }
    int fc = child_lb + 8 * (int)rank[i];
    first_child[node] = fc;
    float x = (float)center[3 * node], y = (float)center[3 * node + 1], z = (float)center[3 * node + 2];
    float w = (float)ph.x, h = (float)ph.y, d = (float)ph.z;
    float w2 = w / 2.0f, h2 = h / 2.0f, d2 = d / 2.0f;
    for (int o = 0; o < 8; o++) {
        int c = fc + o;
        float cx = c_child_sign[o][0] < 0 ? __fsub_rn(x, w2) : __fadd_rn(x, w2);
        float cy = c_child_sign[o][1] < 0 ? __fsub_rn(y, h2) : __fadd_rn(y, h2);
        float cz = c_child_sign[o][2] < 0 ? __fsub_rn(z, d2) : __fadd_rn(z, d2);
        center[3 * c] = cx;
        center[3 * c + 1] = cy;
        center[3 * c + 2] = cz;
        parent[c] = node;
        level[c] = child_level;
        arrivals[c] = 0;
        first_child[c] = -1;
    }
}

// membership of every slot of a subdividing node in its 8 children (usually exactly one)
__global__ void k_child_mask(const double* __restrict__ P, const int* __restrict__ idx, const int* __restrict__ node,
                             int n, const int* __restrict__ first_child, const double* __restrict__ center,
                             double3 ch, uint32_t* __restrict__ count, uint8_t* __restrict__ mask) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int fc = first_child[node[i]];
    uint32_t m = 0;
    if (fc >= 0) {
        int q = idx[i];
        double x = P[3 * q], y = P[3 * q + 1], z = P[3 * q + 2];
        for (int o = 0; o < 8; o++) {
            const double* c = center + 3 * (fc + o);
            if (in_range(x, c[0], ch.x) && in_range(y, c[1], ch.y) && in_range(z, c[2], ch.z)) m |= 1u << o;
        }
    }
    mask[i] = (uint8_t)m;
    count[i] = __popc(m);
}

__global__ void k_emit_children(const int* __restrict__ idx, const int* __restrict__ node, int n,
                                const int* __restrict__ first_child, const uint8_t* __restrict__ mask,
                                const uint32_t* __restrict__ pos, int child_lb, uint32_t* __restrict__ key,
                                int* __restrict__ val) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t m = mask[i];
    if (!m) return;
    int fc = first_child[node[i]] - child_lb;
    uint32_t p = pos[i];
    for (int o = 0; o < 8; o++)
        if (m & (1u << o)) {
            key[p] = (uint32_t)(fc + o);
            val[p] = idx[i];
            p++;
        }
}

// keys are sorted: each child's arrivals are the length of its run of equal keys
__global__ void k_key_runs(const uint32_t* __restrict__ key, int n, int child_lb, int* __restrict__ node,
                           int* __restrict__ arrivals, int* __restrict__ run_start) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = key[i];
    int c = child_lb + (int)k;
    node[i] = c;
    if (i == 0 || key[i - 1] != k) run_start[c] = i;
    if (i == n - 1 || key[i + 1] != k) arrivals[c] = i + 1;   // run end; children are created with 0
}
__global__ void k_run_lengths(int* __restrict__ arrivals, const int* __restrict__ run_start, int lb, int nn) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    int c = lb + i;
    if (arrivals[c] > 0) arrivals[c] -= run_start[c];
}

// bottom-up subtree totals (leaf: arrivals)
__global__ void k_totals(const int* __restrict__ first_child, const int* __restrict__ arrivals, int lb, int nnodes,
                         int* __restrict__ total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    int node = lb + i;
    int fc = first_child[node];
    if (fc < 0) { total[node] = arrivals[node]; return; }
    int s = 0;
    for (int o = 0; o < 8; o++) s += total[fc + o];
    total[node] = s;
}

// top-down DFS ranges
__global__ void k_ranges(const int* __restrict__ first_child, const int* __restrict__ total, int lb, int nnodes,
                         int* __restrict__ begin, int* __restrict__ end) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    int node = lb + i;
    int b = begin[node];
    end[node] = b + total[node];
    int fc = first_child[node];
    if (fc < 0) return;
    for (int o = 0; o < 8; o++) {
        begin[fc + o] = b;
        b += total[fc + o];
    }
}

// leaf slots of one level go to their DFS position
__global__ void k_place_leaves(const int* __restrict__ idx, const int* __restrict__ node, int n,
                               const int* __restrict__ first_child, const int* __restrict__ begin,
                               const uint32_t* __restrict__ seg_start, int lb, int* __restrict__ slots) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int nd = node[i];
    if (first_child[nd] >= 0) return;
    slots[begin[nd] + (i - (int)seg_start[nd - lb])] = idx[i];
}

__global__ void k_leaf_npts(const int* __restrict__ first_child, const int* __restrict__ arrivals, int n,
                            int* __restrict__ npts) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    npts[i] = first_child[i] < 0 ? arrivals[i] : 0;
}

static void grow_keep(DevBuf& b, size_t need, size_t used, hipStream_t st) {
    if (need <= b.bytes && b.ptr) return;
    size_t nb = std::max(need, b.bytes * 2);
    void* np = nullptr;
    FMM_HIP(hipMalloc(&np, nb));
    if (b.ptr && used) FMM_HIP(hipMemcpyAsync(np, b.ptr, used, hipMemcpyDeviceToDevice, st));
    if (b.ptr) {
        FMM_HIP(hipStreamSynchronize(st));
        FMM_HIP(hipFree(b.ptr));
    }
    b.ptr = np;
    b.bytes = nb;
}


static void set_level_sizes(TreeDev& T, const double root_h[3], bool cubic_scale) {
    // per-level half sizes: child half = (float)h / 2 (quadrature_octree.cpp:83-89)
    for (int a = 0; a < 3; a++) T.half[0][a] = root_h[a];
    for (int k = 1; k < MAX_LEVELS; k++)
        for (int a = 0; a < 3; a++) T.half[k][a] = (double)((float)T.half[k - 1][a] / 2.0f);
    for (int k = 0; k < MAX_LEVELS; k++) {
        const double* h = T.half[k];
        T.radius[k] = std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
        // the largest half size keeps every scaled position of a box's points within |u| <= sqrt(3)
        T.scale[k] = cubic_scale ? std::max(h[0], std::max(h[1], h[2])) : T.radius[k] / std::sqrt(3.0);
    }
}

// the host mirrors an evaluation needs (node emptiness and children); the rest of the node
// arrays stay on the device until fetch_tree_details asks for them
static void copy_host_mirrors(TreeDev& T, hipStream_t st) {
    int num_nodes = T.num_nodes;
    T.h_first_child.resize(num_nodes);
    T.h_npts.resize(num_nodes);
    T.details_on_host = false;
    FMM_HIP(hipMemcpyAsync(T.h_first_child.data(), T.first_child.ptr, num_nodes * sizeof(int), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipMemcpyAsync(T.h_npts.data(), T.npts.ptr, num_nodes * sizeof(int), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
}

void fetch_tree_details(TreeDev& T) {
    if (T.details_on_host) return;
    const int num_nodes = T.num_nodes;
    T.h_begin.resize(num_nodes);
    T.h_end.resize(num_nodes);
    T.h_level.resize(num_nodes);
    T.h_parent.resize(num_nodes);
    T.h_center.resize(3 * (size_t)num_nodes);
    FMM_HIP(hipDeviceSynchronize());   // the tree may have been built on either stream
    if (num_nodes) {
        FMM_HIP(hipMemcpy(T.h_begin.data(), T.begin.ptr, num_nodes * sizeof(int), hipMemcpyDeviceToHost));
        FMM_HIP(hipMemcpy(T.h_end.data(), T.end.ptr, num_nodes * sizeof(int), hipMemcpyDeviceToHost));
        FMM_HIP(hipMemcpy(T.h_level.data(), T.level.ptr, num_nodes * sizeof(int), hipMemcpyDeviceToHost));
        FMM_HIP(hipMemcpy(T.h_parent.data(), T.parent.ptr, num_nodes * sizeof(int), hipMemcpyDeviceToHost));
        FMM_HIP(hipMemcpy(T.h_center.data(), T.center.ptr, 3 * (size_t)num_nodes * sizeof(double),
                          hipMemcpyDeviceToHost));
    }
    T.details_on_host = true;
}

// Level-by-level build: every level partitions the points still being refined into the
// children by testing all 8 child boxes, so duplicates (overlaps) and drops (gaps) are exact.
static void build_tree_levels(TreeDev& T, TreeScratch& S, const double* d_xyz, int n, const double root_c[3],
                              int thr, int max_depth, hipStream_t st) {
    // root
    size_t node_cap = 1024;
    grow_keep(T.center, node_cap * 3 * sizeof(double), 0, st);
    grow_keep(T.first_child, node_cap * sizeof(int), 0, st);
    grow_keep(T.parent, node_cap * sizeof(int), 0, st);
    grow_keep(T.level, node_cap * sizeof(int), 0, st);
    grow_keep(T.arrivals, node_cap * sizeof(int), 0, st);
    uint32_t* flag = S.flag.get<uint32_t>(std::max(n, 1));
    uint32_t* pos = S.pos.get<uint32_t>(std::max(n, 1) + 1);
    if (S.lvl_idx.size() < (size_t)MAX_LEVELS) {
        S.lvl_idx.resize(MAX_LEVELS);
        S.lvl_node.resize(MAX_LEVELS);
    }
    S.lvl_n.assign(MAX_LEVELS, 0);
    double3 rc = make_double3(root_c[0], root_c[1], root_c[2]);
    double3 rh = make_double3(T.half[0][0], T.half[0][1], T.half[0][2]);
    int n0 = 0;
    if (n > 0) {
        k_root_filter<<<blocks(n), 256, 0, st>>>(d_xyz, n, rc, rh, flag);
        exclusive_scan_u32(flag, pos, n, S.tmp, st);
        uint32_t last_pos, last_flag;
        FMM_HIP(hipMemcpyAsync(&last_pos, pos + n - 1, 4, hipMemcpyDeviceToHost, st));
        FMM_HIP(hipMemcpyAsync(&last_flag, flag + n - 1, 4, hipMemcpyDeviceToHost, st));
        FMM_HIP(hipStreamSynchronize(st));
        n0 = (int)(last_pos + last_flag);
    }
    T.dropped = n - n0;
    int* idx0 = (int*)S.lvl_idx[0].reserve(std::max(n0, 1) * sizeof(int));
    int* nd0 = (int*)S.lvl_node[0].reserve(std::max(n0, 1) * sizeof(int));
    if (n0 > 0) k_compact_root<<<blocks(n), 256, 0, st>>>(flag, pos, n, idx0, nd0);
    S.lvl_n[0] = n0;
    {
        double c3[3] = {root_c[0], root_c[1], root_c[2]};
        int zero = 0, m1 = -1;
        FMM_HIP(hipMemcpyAsync(T.center.ptr, c3, 3 * sizeof(double), hipMemcpyHostToDevice, st));
        FMM_HIP(hipMemcpyAsync(T.parent.ptr, &m1, sizeof(int), hipMemcpyHostToDevice, st));
        FMM_HIP(hipMemcpyAsync(T.level.ptr, &zero, sizeof(int), hipMemcpyHostToDevice, st));
        FMM_HIP(hipMemcpyAsync(T.first_child.ptr, &m1, sizeof(int), hipMemcpyHostToDevice, st));
        FMM_HIP(hipMemcpyAsync(T.arrivals.ptr, &n0, sizeof(int), hipMemcpyHostToDevice, st));
        FMM_HIP(hipStreamSynchronize(st));
    }
    T.level_begin[0] = 0;
    T.level_begin[1] = 1;
    int num_nodes = 1;
    int k = 0;
    for (;; k++) {
        int lb = T.level_begin[k], le = T.level_begin[k + 1];
        int nn = le - lb;
        uint32_t* sflag = S.sflag.get<uint32_t>(nn + 1);
        uint32_t* srank = S.srank.get<uint32_t>(nn + 1);
        int allow = (k < max_depth) ? 1 : 0;
        k_subdiv_flags<<<blocks(nn), 256, 0, st>>>(T.arrivals.as<int>(), lb, nn, thr, allow, sflag);
        exclusive_scan_u32(sflag, srank, nn, S.tmp, st);
        uint32_t lr, lf;
        FMM_HIP(hipMemcpyAsync(&lr, srank + nn - 1, 4, hipMemcpyDeviceToHost, st));
        FMM_HIP(hipMemcpyAsync(&lf, sflag + nn - 1, 4, hipMemcpyDeviceToHost, st));
        FMM_HIP(hipStreamSynchronize(st));
        int nsub = (int)(lr + lf);
        int child_lb = le;
        int new_nodes = num_nodes + 8 * nsub;
        size_t used = (size_t)num_nodes;
        grow_keep(T.center, (size_t)new_nodes * 3 * sizeof(double), used * 3 * sizeof(double), st);
        grow_keep(T.first_child, (size_t)new_nodes * sizeof(int), used * sizeof(int), st);
        grow_keep(T.parent, (size_t)new_nodes * sizeof(int), used * sizeof(int), st);
        grow_keep(T.level, (size_t)new_nodes * sizeof(int), used * sizeof(int), st);
        grow_keep(T.arrivals, (size_t)new_nodes * sizeof(int), used * sizeof(int), st);
        double3 ph = make_double3(T.half[k][0], T.half[k][1], T.half[k][2]);
        k_make_children<<<blocks(nn), 256, 0, st>>>(sflag, srank, lb, nn, child_lb, k + 1, ph,
                                                    T.center.as<double>(), T.first_child.as<int>(),
                                                    T.parent.as<int>(), T.level.as<int>(), T.arrivals.as<int>());
        num_nodes = new_nodes;
        T.level_begin[k + 2] = new_nodes;
        if (nsub == 0) break;
        // partition the level's slots into the children
        int nk = S.lvl_n[k];
        uint32_t* cnt = S.flag.get<uint32_t>(std::max(nk, 1) + 1);
        uint32_t* off = S.pos.get<uint32_t>(std::max(nk, 1) + 1);
        uint8_t* mask = S.mask.get<uint8_t>(std::max(nk, 1));
        double3 chh = make_double3(T.half[k + 1][0], T.half[k + 1][1], T.half[k + 1][2]);
        const int* idxk = (const int*)S.lvl_idx[k].ptr;
        const int* ndk = (const int*)S.lvl_node[k].ptr;
        k_child_mask<<<blocks(nk), 256, 0, st>>>(d_xyz, idxk, ndk, nk, T.first_child.as<int>(),
                                                 T.center.as<double>(), chh, cnt, mask);
        exclusive_scan_u32(cnt, off, nk, S.tmp, st);
        uint32_t lo, lc;
        FMM_HIP(hipMemcpyAsync(&lo, off + nk - 1, 4, hipMemcpyDeviceToHost, st));
        FMM_HIP(hipMemcpyAsync(&lc, cnt + nk - 1, 4, hipMemcpyDeviceToHost, st));
        FMM_HIP(hipStreamSynchronize(st));
        int nk1 = (int)(lo + lc);
        uint32_t* key = S.key.get<uint32_t>(std::max(nk1, 1));
        int* val = S.val.get<int>(std::max(nk1, 1));
        uint32_t* key2 = S.key2.get<uint32_t>(std::max(nk1, 1));
        int* idx1 = (int*)S.lvl_idx[k + 1].reserve(std::max(nk1, 1) * sizeof(int));
        int* nd1 = (int*)S.lvl_node[k + 1].reserve(std::max(nk1, 1) * sizeof(int));
        k_emit_children<<<blocks(nk), 256, 0, st>>>(idxk, ndk, nk, T.first_child.as<int>(), mask, off, child_lb, key,
                                                    val);
        int nchild = 8 * nsub;
        int bits = 1;
        while ((1 << bits) < nchild) bits++;
        sort_pairs_u32(key, key2, val, idx1, nk1, bits, S.tmp, st);
        int* rs = S.total.get<int>(new_nodes);
        k_key_runs<<<blocks(nk1), 256, 0, st>>>(key2, nk1, child_lb, nd1, T.arrivals.as<int>(), rs);
        k_run_lengths<<<blocks(nchild), 256, 0, st>>>(T.arrivals.as<int>(), rs, child_lb, nchild);
        S.lvl_n[k + 1] = nk1;
    }
    T.num_levels = k + 1;   // levels 0..k hold nodes
    T.num_nodes = num_nodes;
    // totals bottom-up, ranges top-down
    T.npts.reserve((size_t)num_nodes * sizeof(int));
    T.begin.reserve((size_t)num_nodes * sizeof(int));
    T.end.reserve((size_t)num_nodes * sizeof(int));
    int* total = S.total.get<int>(num_nodes);
    for (int lv = T.num_levels; lv >= 0; lv--) {
        int lb = T.level_begin[lv], nn = T.level_begin[lv + 1] - lb;
        if (nn > 0) k_totals<<<blocks(nn), 256, 0, st>>>(T.first_child.as<int>(), T.arrivals.as<int>(), lb, nn, total);
    }
    {
        int zero = 0;
        FMM_HIP(hipMemcpyAsync(T.begin.ptr, &zero, sizeof(int), hipMemcpyHostToDevice, st));
    }
    for (int lv = 0; lv <= T.num_levels; lv++) {
        int lb = T.level_begin[lv], nn = T.level_begin[lv + 1] - lb;
        if (nn > 0) k_ranges<<<blocks(nn), 256, 0, st>>>(T.first_child.as<int>(), total, lb, nn, T.begin.as<int>(),
                                                       T.end.as<int>());
    }
    int nslots = 0;
    FMM_HIP(hipMemcpyAsync(&nslots, total, sizeof(int), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
    T.num_slots = nslots;
    int* slots = T.slots.get<int>(std::max(nslots, 1));
    for (int lv = 0; lv <= k; lv++) {
        int nk = S.lvl_n[lv];
        if (nk == 0) continue;
        int lb = T.level_begin[lv], nn = T.level_begin[lv + 1] - lb;
        // segment start of each level node inside the level arrays = prefix of arrivals
        uint32_t* seg = S.srank.get<uint32_t>(nn + 1);
        exclusive_scan_u32((const uint32_t*)(T.arrivals.as<int>() + lb), seg, nn, S.tmp, st);
        k_place_leaves<<<blocks(nk), 256, 0, st>>>((const int*)S.lvl_idx[lv].ptr, (const int*)S.lvl_node[lv].ptr, nk,
                                                  T.first_child.as<int>(), T.begin.as<int>(), seg, lb, slots);
    }
    k_leaf_npts<<<blocks(num_nodes), 256, 0, st>>>(T.first_child.as<int>(), T.arrivals.as<int>(), num_nodes,
                                                   T.npts.as<int>());
}

// ---------------------------------------------------------------- one-sort build
// Each point's path (octant digit per level, reference child order) is computed once from
// the same float-rounded child boxes and half-open tests; one radix sort of the root's
// points by path puts them in DFS order, so every node is a contiguous range of the sorted
// array.  Levels are then built top-down point-parallel: each sorted point carries its
// node, moves to child (first_child + digit) when the node subdivides, and a child's range
// starts/ends where the neighbour's path prefix differs.  A point whose child at some level
// is not unique (a gap or an overlap of the float-rounded boxes) is marked with that level;
// if such a level is reached by the point's leaf the tree is rebuilt level by level, which
// keeps duplicates and drops exact.
constexpr int PATH_LEVELS = 21;   // 3 bits per level in a 63-bit key (level 1 in the top bits)
constexpr int SHORT_LEVELS = 10;  // first attempt: 30-bit keys

struct LevelHalves {
    float h[PATH_LEVELS + 1][3];   // half size of level k boxes (float-rounded, as the tree's)
};

__device__ __forceinline__ int octant_of(bool sx, bool sy, bool sz) {
    // inverse of CHILD_SIGN: (-,-)->0, (+,-)->1, (+,+)->2, (-,+)->3, +4 for +z
    return (sz ? 4 : 0) + (sx ? (sy ? 2 : 1) : (sy ? 3 : 0));
}

__global__ void k_paths(const double* __restrict__ P, int n, double3 rc, double3 rh, LevelHalves H, int depth,
                        uint64_t* __restrict__ key, int* __restrict__ irregular, uint32_t* __restrict__ inroot) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = P[3 * i], y = P[3 * i + 1], z = P[3 * i + 2];
    bool in = in_range(x, rc.x, rh.x) && in_range(y, rc.y, rh.y) && in_range(z, rc.z, rh.z);
    inroot[i] = in ? 1u : 0u;
    if (!in) {
        key[i] = 0;
        irregular[i] = 0;
        return;
    }
    float cx = (float)rc.x, cy = (float)rc.y, cz = (float)rc.z;
    uint64_t k = 0;
    int irr = PATH_LEVELS + 1;
    for (int l = 1; l <= depth; l++) {
        const float hx = H.h[l][0], hy = H.h[l][1], hz = H.h[l][2];
        const float lx = __fsub_rn(cx, hx), ux = __fadd_rn(cx, hx);
        const float ly = __fsub_rn(cy, hy), uy = __fadd_rn(cy, hy);
        const float lz = __fsub_rn(cz, hz), uz = __fadd_rn(cz, hz);
        const double dhx = hx, dhy = hy, dhz = hz;
        const bool x0 = in_range(x, lx, dhx), x1 = in_range(x, ux, dhx);
        const bool y0 = in_range(y, ly, dhy), y1 = in_range(y, uy, dhy);
        const bool z0 = in_range(z, lz, dhz), z1 = in_range(z, uz, dhz);
        if (x0 == x1 || y0 == y1 || z0 == z1) {   // in two children or in none
            irr = l;
            k <<= 3 * (depth - l + 1);
            break;
        }
        k = (k << 3) | (uint64_t)octant_of(x1, y1, z1);
        cx = x1 ? ux : lx;
        cy = y1 ? uy : ly;
        cz = z1 ? uz : lz;
    }
    key[i] = k;
    irregular[i] = irr;
}

// the root's points, compacted in input order, keyed by the top `levels` levels of their path
template <class K>
__global__ void k_compact_paths(const uint64_t* __restrict__ key, const uint32_t* __restrict__ inroot,
                                const uint32_t* __restrict__ rank, int n, int shift, K* __restrict__ kc,
                                int* __restrict__ idx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !inroot[i]) return;
    kc[rank[i]] = (K)(key[i] >> shift);
    idx[rank[i]] = i;
}

__global__ void k_sorted_flags(const int* __restrict__ sb, const int* __restrict__ se, int lb, int nnodes, int thr,
                               int allow, int deeper, uint32_t* __restrict__ flag, int* __restrict__ need_deeper) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    int node = lb + i;
    bool sub = se[node] - sb[node] >= thr;
    flag[i] = (allow && sub) ? 1u : 0u;
    if (deeper && sub) *need_deeper = 1;
}

// number of subdividing nodes and the need-deeper flag in one readback
__global__ void k_level_counts(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ rank, int nn,
                               const int* __restrict__ need, int* __restrict__ out) {
    out[0] = (int)(rank[nn - 1] + flag[nn - 1]);
    out[1] = *need;
}

// the 8 children of every subdivided node: float centres as the reference; ranges start empty
__global__ void k_make_children_sorted(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ rank, int lb,
                                       int nnodes, int child_lb, int child_level, double3 ph,
                                       double* __restrict__ center, int* __restrict__ first_child,
                                       int* __restrict__ parent, int* __restrict__ level, int* __restrict__ sb,
                                       int* __restrict__ se) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    int node = lb + i;
    if (!flag[i]) { first_child[node] = -1; return; }
    int fc = child_lb + 8 * (int)rank[i];
    first_child[node] = fc;
    float x = (float)center[3 * node], y = (float)center[3 * node + 1], z = (float)center[3 * node + 2];
    float w = (float)ph.x, h = (float)ph.y, d = (float)ph.z;
    float w2 = w / 2.0f, h2 = h / 2.0f, d2 = d / 2.0f;
    for (int o = 0; o < 8; o++) {
        int c = fc + o;
        center[3 * c] = c_child_sign[o][0] < 0 ? __fsub_rn(x, w2) : __fadd_rn(x, w2);
        center[3 * c + 1] = c_child_sign[o][1] < 0 ? __fsub_rn(y, h2) : __fadd_rn(y, h2);
        center[3 * c + 2] = c_child_sign[o][2] < 0 ? __fsub_rn(z, d2) : __fadd_rn(z, d2);
        parent[c] = node;
        level[c] = child_level;
        first_child[c] = -1;
        sb[c] = -1;
        se[c] = -1;
    }
}

// every sorted point of a subdivided node moves to its child; runs of equal path prefix
// give the children's ranges
template <class K>
__global__ void k_assign_level(const K* __restrict__ sk, int n0, int child_level, int depth,
                               const int* __restrict__ first_child, int* __restrict__ node_of, int* __restrict__ sb,
                               int* __restrict__ se) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n0) return;
    int fc = first_child[node_of[j]];
    if (fc < 0) return;
    const int sh = 3 * (depth - child_level);
    const K pre = sk[j] >> sh;
    const int c = fc + (int)(pre & (K)7);
    node_of[j] = c;
    if (j == 0 || (sk[j - 1] >> sh) != pre) sb[c] = j;
    if (j == n0 - 1 || (sk[j + 1] >> sh) != pre) se[c] = j + 1;
}

// empty children: begin = end = where they sit between their siblings (DFS position)
__global__ void k_fix_empty(const int* __restrict__ first_child, int lb, int nnodes, int* __restrict__ sb,
                            int* __restrict__ se) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    int node = lb + i;
    int fc = first_child[node];
    if (fc < 0) return;
    int pos = sb[node];
    for (int o = 0; o < 8; o++) {
        int c = fc + o;
        if (sb[c] < 0) sb[c] = se[c] = pos;
        else pos = se[c];
    }
}

__global__ void k_node_ranges(const int* __restrict__ first_child, const int* __restrict__ sb,
                              const int* __restrict__ se, int nnodes, int* __restrict__ npts, int* __restrict__ begin,
                              int* __restrict__ end) {
    int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nnodes) return;
    begin[node] = sb[node];
    end[node] = se[node];
    npts[node] = first_child[node] < 0 ? se[node] - sb[node] : 0;
}

// per sorted point: the key of its leaf for the in-leaf (input order) sort, and the check
// that no point's path was ambiguous at or above its leaf
__global__ void k_leaf_keys(const int* __restrict__ node_of, const int* __restrict__ order, int n0,
                            const uint32_t* __restrict__ rank, const int* __restrict__ sb, const int* __restrict__ lvl,
                            const int* __restrict__ irregular, uint32_t* __restrict__ leaf_key, int* __restrict__ bad) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n0) return;
    int nd = node_of[j], q = order[j];
    leaf_key[rank[q]] = (uint32_t)sb[nd];
    if (irregular[q] <= lvl[nd]) *bad = 1;
}

// In-leaf input order without a second sort: point j (path-sorted position) of leaf nd, whose
// points are the positions [sb, se), goes to slot sb + (number of the leaf's points with a
// smaller input index).  flags[0]: a point's ambiguous level is at or above its leaf (the tree
// must be rebuilt level by level); flags[1]: a leaf holds more than LEAF_RANK_MAX points (the
// quadratic count is skipped and the stable sort by leaf start orders the slots instead).
constexpr int LEAF_RANK_MAX = 512;
__global__ void k_leaf_slots(const int* __restrict__ node_of, const int* __restrict__ order, int n0,
                             const int* __restrict__ sb, const int* __restrict__ se, const int* __restrict__ lvl,
                             const int* __restrict__ irregular, int* __restrict__ slots, int* __restrict__ flags) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n0) return;
    const int nd = node_of[j], q = order[j];
    if (irregular[q] <= lvl[nd]) flags[0] = 1;
    const int b = sb[nd], e = se[nd];
    if (e - b > LEAF_RANK_MAX) {
        flags[1] = 1;
        return;
    }
    int r = 0;
    for (int u = b; u < e; u++) r += order[u] < q;
    slots[b + r] = q;
}

__global__ void k_count_root(const uint32_t* __restrict__ inroot, const uint32_t* __restrict__ rank, int n,
                             int* __restrict__ out) {
    out[0] = (int)(rank[n - 1] + inroot[n - 1]);
}

__global__ void k_init_root(double3 c, int n0, double* __restrict__ center, int* __restrict__ parent,
                            int* __restrict__ level, int* __restrict__ first_child, int* __restrict__ sb,
                            int* __restrict__ se, int* __restrict__ flags2) {
    center[0] = c.x;
    center[1] = c.y;
    center[2] = c.z;
    parent[0] = -1;
    level[0] = 0;
    first_child[0] = -1;
    sb[0] = 0;
    se[0] = n0;
    flags2[0] = flags2[1] = 0;
}
__global__ void k_zero_i32(int* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// A level with its boundaries kept on the device (no host sync): k_level_children numbers the
// subdividing nodes of level k (their per-chunk counts come from the level before, chunk prefix
// + block scan, i.e. parent order) and makes their 8 children (float centres as the
// reference, empty sorted ranges); k_assign_count moves the points to the new children, gives
// the children's ranges and counts the subdividing children per chunk of their level.  The
// chunks are LEVEL_WGS contiguous pieces of the nodes [level_begin[k], level_begin[k+1]).
// Children beyond node_cap set flags[1] and none are made; a node below the key's depth that
// would subdivide sets flags[0].
constexpr int LEVEL_WGS = 256;

__device__ __forceinline__ int block_excl_sum(int v, int& total) {   // 256 threads
    __shared__ int wsum[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = v;
    for (int d = 1; d < 64; d <<= 1) {
        int u = __shfl_up(incl, d, 64);
        if (lane >= d) incl += u;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int off = 0;
    for (int q = 0; q < w; q++) off += wsum[q];
    total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return off + incl - v;
}

__device__ __forceinline__ void level_chunk(const int* level_begin, int k, int& a, int& b) {
    const int lb = level_begin[k], le = level_begin[k + 1], nn = le - lb;
    const int per_wg = (nn + LEVEL_WGS - 1) / LEVEL_WGS;
    const int c0 = min(nn, blockIdx.x * per_wg), c1 = min(nn, c0 + per_wg);
    const int per_t = (c1 - c0 + 255) / 256;
    a = lb + min(c1, c0 + (int)threadIdx.x * per_t);
    b = lb + min(c1, c0 + ((int)threadIdx.x + 1) * per_t);
}

// level k may subdivide its nodes / a level-k node that would subdivide needs a longer key
__device__ __forceinline__ bool level_allows(int k, int depth, int max_depth) { return k < max_depth && k < depth; }
__device__ __forceinline__ bool level_deeper(int k, int depth, int max_depth) { return k < max_depth && k >= depth; }

// part: subdividing nodes per chunk of level k (filled by the level before); next: the same
// for level k + 1, cleared here for k_assign_count
__global__ __launch_bounds__(256) void k_level_children(int k, int depth, int max_depth, int thr, double3 ph,
                                                        int node_cap, int* __restrict__ level_begin,
                                                        int* __restrict__ flags, const int* __restrict__ part,
                                                        int* __restrict__ next, double* __restrict__ center,
                                                        int* __restrict__ first_child, int* __restrict__ parent,
                                                        int* __restrict__ level, int* __restrict__ sb,
                                                        int* __restrict__ se) {
    const bool allow = level_allows(k, depth, max_depth);
    const int le = level_begin[k + 1];
    // subdividing nodes before this chunk, and in the whole level
    int before = 0, all = 0;
    for (int q = threadIdx.x; q < LEVEL_WGS; q += 256) {
        int v = part[q];
        all += v;
        if (q < (int)blockIdx.x) before += v;
    }
    int tb, ta;
    block_excl_sum(before, tb);
    block_excl_sum(all, ta);
    if (threadIdx.x == 0) next[blockIdx.x] = 0;
    const bool fits = (long long)le + 8LL * ta <= (long long)node_cap;
    int a, b;
    level_chunk(level_begin, k, a, b);
    int cnt = 0;
    for (int i = a; i < b; i++) cnt += (allow && se[i] - sb[i] >= thr) ? 1 : 0;
    int chunk_total;
    int off = tb + block_excl_sum(cnt, chunk_total);
    float w2 = (float)ph.x / 2.0f, h2 = (float)ph.y / 2.0f, d2 = (float)ph.z / 2.0f;
    for (int i = a; i < b; i++) {
        bool sub = allow && fits && se[i] - sb[i] >= thr;
        if (!sub) { first_child[i] = -1; continue; }
        int fc = le + 8 * off++;
        first_child[i] = fc;
        float x = (float)center[3 * i], y = (float)center[3 * i + 1], z = (float)center[3 * i + 2];
        for (int o = 0; o < 8; o++) {
            int c = fc + o;
            center[3 * c] = c_child_sign[o][0] < 0 ? __fsub_rn(x, w2) : __fadd_rn(x, w2);
            center[3 * c + 1] = c_child_sign[o][1] < 0 ? __fsub_rn(y, h2) : __fadd_rn(y, h2);
            center[3 * c + 2] = c_child_sign[o][2] < 0 ? __fsub_rn(z, d2) : __fadd_rn(z, d2);
            parent[c] = i;
            level[c] = k + 1;
            first_child[c] = -1;
            sb[c] = -1;
            se[c] = -1;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (!fits) flags[1] = 1;
        level_begin[k + 2] = le + (fits ? 8 * ta : 0);
    }
}

// every sorted point of a subdivided level-k node moves to its child; runs of equal path
// prefix give the children's ranges.  The first point of a child's run tests whether the run
// reaches thr points (the keys are sorted, so it does iff point j + thr - 1 has the same
// prefix) and counts the child into its chunk of level k + 1.
template <class K>
__global__ void k_assign_count(const K* __restrict__ sk, int n0, int k, int depth, int max_depth, int thr,
                               const int* __restrict__ level_begin, const int* __restrict__ first_child,
                               int* __restrict__ node_of, int* __restrict__ sb, int* __restrict__ se,
                               int* __restrict__ next, int* __restrict__ flags) {
    const int lb1 = level_begin[k + 1], le1 = level_begin[k + 2];
    if (le1 == lb1) return;   // no node of level k subdivided
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n0) return;
    const int fc = first_child[node_of[j]];
    if (fc < 0) return;
    const int child_level = k + 1, sh = 3 * (depth - child_level);
    const K pre = sk[j] >> sh;
    const int c = fc + (int)(pre & (K)7);
    node_of[j] = c;
    if (j == 0 || (sk[j - 1] >> sh) != pre) {
        sb[c] = j;
        const bool sub = j + thr - 1 < n0 && (sk[j + thr - 1] >> sh) == pre;
        if (sub && level_deeper(child_level, depth, max_depth)) flags[0] = 1;
        if (sub && level_allows(child_level, depth, max_depth)) {
            const int per_wg = (le1 - lb1 + LEVEL_WGS - 1) / LEVEL_WGS;
            atomicAdd(&next[(c - lb1) / per_wg], 1);
        }
    }
    if (j == n0 - 1 || (sk[j + 1] >> sh) != pre) se[c] = j + 1;
}

// empty children of every subdivided node, all levels at once (a node's own range is final
// once the points are assigned; empty nodes never subdivide): begin = end = where they sit
// between their siblings (DFS position)
__global__ void k_fix_empty_all(const int* __restrict__ first_child, const int* __restrict__ level_begin, int total_at,
                                int* __restrict__ sb, int* __restrict__ se) {
    const int nn = level_begin[total_at];
    for (int node = blockIdx.x * blockDim.x + threadIdx.x; node < nn; node += gridDim.x * blockDim.x) {
        int fc = first_child[node];
        if (fc < 0) continue;
        int pos = sb[node];
        for (int o = 0; o < 8; o++) {
            int c = fc + o;
            if (sb[c] < 0) sb[c] = se[c] = pos;
            else pos = se[c];
        }
    }
}

// the root, the level boundaries, the flags and the chunk counts of level 0 (one workgroup)
__global__ void k_init_root_dev(double3 c, int n0, int thr, int depth, int max_depth, double* __restrict__ center,
                                int* __restrict__ parent, int* __restrict__ level, int* __restrict__ first_child,
                                int* __restrict__ sb, int* __restrict__ se, int* __restrict__ level_begin, int nlb,
                                int* __restrict__ flags, int* __restrict__ part) {
    for (int q = threadIdx.x; q < 2 * LEVEL_WGS; q += blockDim.x) part[q] = 0;
    if (threadIdx.x != 0) return;
    center[0] = c.x;
    center[1] = c.y;
    center[2] = c.z;
    parent[0] = -1;
    level[0] = 0;
    first_child[0] = -1;
    sb[0] = 0;
    se[0] = n0;
    level_begin[0] = 0;
    for (int i = 1; i < nlb; i++) level_begin[i] = 1;
    const bool sub = n0 >= thr;
    flags[0] = (sub && level_deeper(0, depth, max_depth)) ? 1 : 0;
    flags[1] = 0;
    part[0] = (sub && level_allows(0, depth, max_depth)) ? 1 : 0;   // the root is chunk 0 of level 0
}

// Same tree as build_from_sorted, all levels queued at once: the node arrays are sized for
// the bound (each subdividing node holds >= thr points, so a level adds <= 8 n0 / thr
// nodes) and read back once.  Returns 0 on success, 1 if the key is too short (a node below
// `depth` subdivides), 2 if the node bound was exceeded.
template <class K>
static int build_from_sorted_queued(TreeDev& T, TreeScratch& S, const K* sk, int n0, int depth, int thr,
                                    int max_depth, const double root_c[3], size_t node_cap, hipStream_t st) {
    grow_keep(T.center, node_cap * 3 * sizeof(double), 0, st);
    grow_keep(T.first_child, node_cap * sizeof(int), 0, st);
    grow_keep(T.parent, node_cap * sizeof(int), 0, st);
    grow_keep(T.level, node_cap * sizeof(int), 0, st);
    grow_keep(S.sb, node_cap * sizeof(int), 0, st);
    grow_keep(S.se, node_cap * sizeof(int), 0, st);
    const int nlb = MAX_LEVELS + 2;
    int* dlb = S.dlevel.get<int>(nlb + 2);
    int* flags = dlb + nlb;
    int* node_of = S.node_of.get<int>(std::max(n0, 1));
    int* part = S.partial.get<int>(2 * LEVEL_WGS);   // chunk counts of two consecutive levels
    k_init_root_dev<<<1, 256, 0, st>>>(make_double3(root_c[0], root_c[1], root_c[2]), n0, thr, depth, max_depth,
                                       T.center.as<double>(), T.parent.as<int>(), T.level.as<int>(),
                                       T.first_child.as<int>(), S.sb.as<int>(), S.se.as<int>(), dlb, nlb, flags, part);
    if (n0) k_zero_i32<<<blocks(n0), 256, 0, st>>>(node_of, n0);
    const int last = std::min(max_depth, depth);   // levels 0..last can hold nodes
    for (int k = 0; k <= last; k++) {
        double3 ph = make_double3(T.half[k][0], T.half[k][1], T.half[k][2]);
        int* cur = part + (k & 1) * LEVEL_WGS;
        int* next = part + ((k + 1) & 1) * LEVEL_WGS;
        k_level_children<<<LEVEL_WGS, 256, 0, st>>>(k, depth, max_depth, thr, ph, (int)node_cap, dlb, flags, cur, next,
                                                    T.center.as<double>(), T.first_child.as<int>(),
                                                    T.parent.as<int>(), T.level.as<int>(), S.sb.as<int>(),
                                                    S.se.as<int>());
        if (k == last) break;
        if (n0)
            k_assign_count<K><<<blocks(n0), 256, 0, st>>>(sk, n0, k, depth, max_depth, thr, dlb, T.first_child.as<int>(),
                                                          node_of, S.sb.as<int>(), S.se.as<int>(), next, flags);
    }
    k_fix_empty_all<<<std::max(1, std::min(blocks(node_cap), 4 * num_cus())), 256, 0, st>>>(
        T.first_child.as<int>(), dlb, last + 2, S.sb.as<int>(), S.se.as<int>());
    std::vector<int> h(nlb + 2);
    FMM_HIP(hipMemcpyAsync(h.data(), dlb, (nlb + 2) * sizeof(int), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
    if (h[nlb]) return 1;
    if (h[nlb + 1]) return 2;
    int k = 0;
    while (k + 2 < nlb && h[k + 2] > h[k + 1]) k++;   // level k+1 holds nodes while its range is non-empty
    T.num_levels = k + 1;
    T.num_nodes = h[k + 1];
    for (int i = 0; i < nlb; i++) T.level_begin[i] = std::min(h[i], T.num_nodes);
    return 0;
}

// returns false when the key is too short for the tree (a node below `depth` subdivides)
template <class K>
static bool build_from_sorted(TreeDev& T, TreeScratch& S, const K* sk, int n0, int depth, int thr, int max_depth,
                              const double root_c[3], hipStream_t st) {
    size_t node_cap = 1024;
    grow_keep(T.center, node_cap * 3 * sizeof(double), 0, st);
    grow_keep(T.first_child, node_cap * sizeof(int), 0, st);
    grow_keep(T.parent, node_cap * sizeof(int), 0, st);
    grow_keep(T.level, node_cap * sizeof(int), 0, st);
    grow_keep(S.sb, node_cap * sizeof(int), 0, st);
    grow_keep(S.se, node_cap * sizeof(int), 0, st);
    int* need = S.need.get<int>(4);
    int* node_of = S.node_of.get<int>(std::max(n0, 1));
    k_init_root<<<1, 1, 0, st>>>(make_double3(root_c[0], root_c[1], root_c[2]), n0, T.center.as<double>(),
                                 T.parent.as<int>(), T.level.as<int>(), T.first_child.as<int>(), S.sb.as<int>(),
                                 S.se.as<int>(), need);
    if (n0) k_zero_i32<<<blocks(n0), 256, 0, st>>>(node_of, n0);
    T.level_begin[0] = 0;
    T.level_begin[1] = 1;
    int num_nodes = 1, k = 0;
    for (;; k++) {
        int lb = T.level_begin[k], le = T.level_begin[k + 1], nn = le - lb;
        uint32_t* sflag = S.sflag.get<uint32_t>(nn + 1);
        uint32_t* srank = S.srank.get<uint32_t>(nn + 1);
        // the reference's rule (arrivals >= thr) below max_depth; a node that would subdivide
        // below the key's depth needs the longer key
        int allow = (k < max_depth && k < depth) ? 1 : 0;
        int deeper = (k < max_depth && k >= depth) ? 1 : 0;
        k_sorted_flags<<<blocks(nn), 256, 0, st>>>(S.sb.as<int>(), S.se.as<int>(), lb, nn, thr, allow, deeper, sflag,
                                                   need);
        exclusive_scan_u32(sflag, srank, nn, S.tmp, st);
        k_level_counts<<<1, 1, 0, st>>>(sflag, srank, nn, need, need + 2);
        FMM_HIP(hipMemcpyAsync(S.h_last, need + 2, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        FMM_HIP(hipStreamSynchronize(st));
        if (S.h_last[1]) return false;
        int nsub = (int)S.h_last[0];
        int child_lb = le;
        int new_nodes = num_nodes + 8 * nsub;
        size_t used = (size_t)num_nodes;
        grow_keep(T.center, (size_t)new_nodes * 3 * sizeof(double), used * 3 * sizeof(double), st);
        grow_keep(T.first_child, (size_t)new_nodes * sizeof(int), used * sizeof(int), st);
        grow_keep(T.parent, (size_t)new_nodes * sizeof(int), used * sizeof(int), st);
        grow_keep(T.level, (size_t)new_nodes * sizeof(int), used * sizeof(int), st);
        grow_keep(S.sb, (size_t)new_nodes * sizeof(int), used * sizeof(int), st);
        grow_keep(S.se, (size_t)new_nodes * sizeof(int), used * sizeof(int), st);
        double3 ph = make_double3(T.half[k][0], T.half[k][1], T.half[k][2]);
        k_make_children_sorted<<<blocks(nn), 256, 0, st>>>(sflag, srank, lb, nn, child_lb, k + 1, ph,
                                                           T.center.as<double>(), T.first_child.as<int>(),
                                                           T.parent.as<int>(), T.level.as<int>(), S.sb.as<int>(),
                                                           S.se.as<int>());
        num_nodes = new_nodes;
        T.level_begin[k + 2] = new_nodes;
        if (nsub == 0) break;
        if (n0) k_assign_level<K><<<blocks(n0), 256, 0, st>>>(sk, n0, k + 1, depth, T.first_child.as<int>(), node_of,
                                                               S.sb.as<int>(), S.se.as<int>());
        k_fix_empty<<<blocks(nn), 256, 0, st>>>(T.first_child.as<int>(), lb, nn, S.sb.as<int>(), S.se.as<int>());
    }
    T.num_levels = k + 1;
    T.num_nodes = num_nodes;
    return true;
}

void build_tree(TreeDev& T, TreeScratch& S, const double* d_xyz, int n, const double root_c[3],
                const double root_h[3], int thr, int max_depth, bool cubic_scale, hipStream_t st) {
    if (thr < 2) throw Error(1, "leaf capacity too small: a node would subdivide forever");
    if (max_depth < 0 || max_depth >= MAX_LEVELS) max_depth = MAX_LEVELS - 1;
    set_level_sizes(T, root_h, cubic_scale);
    T.sorted_build = false;
    if (n > 0 && max_depth <= PATH_LEVELS) {
        const int depth = max_depth;
        LevelHalves H;
        for (int l = 0; l <= PATH_LEVELS; l++)
            for (int a = 0; a < 3; a++) H.h[l][a] = (float)T.half[std::min(l, MAX_LEVELS - 1)][a];
        uint64_t* key = S.k64.get<uint64_t>(n);
        int* irr = S.irr.get<int>(n);
        uint32_t* inroot = S.flag.get<uint32_t>(n + 1);
        uint32_t* rank = S.pos.get<uint32_t>(n + 1);
        double3 rc = make_double3(root_c[0], root_c[1], root_c[2]);
        double3 rh = make_double3(root_h[0], root_h[1], root_h[2]);
        // the first attempt needs the top short_depth levels of the paths only; the 64-bit one
        // below recomputes them to full depth
        const int short_depth = std::min(depth, SHORT_LEVELS);
        k_paths<<<blocks(n), 256, 0, st>>>(d_xyz, n, rc, rh, H, short_depth, key, irr, inroot);
        exclusive_scan_u32(inroot, rank, n, S.tmp, st);
        int* cnt2 = S.need.get<int>(4);
        k_count_root<<<1, 1, 0, st>>>(inroot, rank, n, cnt2);
        FMM_HIP(hipMemcpyAsync(S.h_last, cnt2, sizeof(int), hipMemcpyDeviceToHost, st));
        FMM_HIP(hipStreamSynchronize(st));
        const int n0 = (int)S.h_last[0];
        int* idx = S.val.get<int>(std::max(n0, 1));       // root's points in input order
        int* order = S.order.get<int>(std::max(n0, 1));   // ... sorted by path
        bool ok = false;
        if (n0 > 0) {
            uint32_t* k32 = S.key.get<uint32_t>(n0);
            uint32_t* k32s = S.key2.get<uint32_t>(n0);
            k_compact_paths<uint32_t><<<blocks(n), 256, 0, st>>>(key, inroot, rank, n, 0, k32, idx);
            sort_points_u32(k32, k32s, idx, order, n0, 3 * short_depth, S.tmp, st);
            // node bound of the queued build: 8 n0 / thr children per level
            const size_t per_level = 8 * ((size_t)n0 / thr + 1);
            const size_t cap_short = 1 + per_level * (size_t)short_depth;
            const bool queued = cap_short * 44 <= ((size_t)1 << 31);
            int r = queued ? build_from_sorted_queued<uint32_t>(T, S, k32s, n0, short_depth, thr, max_depth, root_c,
                                                                cap_short, st)
                           : (build_from_sorted<uint32_t>(T, S, k32s, n0, short_depth, thr, max_depth, root_c, st) ? 0 : 1);
            if (r == 2) r = build_from_sorted<uint32_t>(T, S, k32s, n0, short_depth, thr, max_depth, root_c, st) ? 0 : 1;
            ok = r == 0;
            if (!ok) {
                uint64_t* kc = S.k64c.get<uint64_t>(n0);
                uint64_t* kcs = S.k64s.get<uint64_t>(n0);
                if (depth > short_depth)
                    k_paths<<<blocks(n), 256, 0, st>>>(d_xyz, n, rc, rh, H, depth, key, irr, inroot);
                k_compact_paths<uint64_t><<<blocks(n), 256, 0, st>>>(key, inroot, rank, n, 0, kc, idx);
                sort_pairs_u64(kc, kcs, idx, order, n0, 3 * depth, S.tmp, st);
                ok = build_from_sorted<uint64_t>(T, S, kcs, n0, depth, thr, max_depth, root_c, st);
            }
        } else {
            ok = build_from_sorted<uint32_t>(T, S, (const uint32_t*)nullptr, 0, short_depth, thr, max_depth, root_c, st);
        }
        if (ok) {
            const int nn = T.num_nodes;
            T.npts.reserve((size_t)nn * sizeof(int));
            T.begin.reserve((size_t)nn * sizeof(int));
            T.end.reserve((size_t)nn * sizeof(int));
            k_node_ranges<<<blocks(nn), 256, 0, st>>>(T.first_child.as<int>(), S.sb.as<int>(), S.se.as<int>(), nn,
                                                      T.npts.as<int>(), T.begin.as<int>(), T.end.as<int>());
            // in-leaf input order: each leaf's points ranked by input index (a stable sort of
            // the input-ordered root points by leaf start if a leaf is too large for that)
            int* slots = T.slots.get<int>(std::max(n0, 1));
            int* flags = S.need.get<int>(4) + 2;
            FMM_HIP(hipMemsetAsync(flags, 0, 2 * sizeof(int), st));
            if (n0) k_leaf_slots<<<blocks(n0), 256, 0, st>>>(S.node_of.as<int>(), order, n0, S.sb.as<int>(),
                                                             S.se.as<int>(), T.level.as<int>(), irr, slots, flags);
            FMM_HIP(hipMemcpyAsync(S.h_last, flags, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
            FMM_HIP(hipStreamSynchronize(st));
            if (S.h_last[0] == 0) {
                if (S.h_last[1] && n0) {
                    uint32_t* lkey = S.key.get<uint32_t>(n0);
                    uint32_t* lkeys = S.key2.get<uint32_t>(n0);
                    int* bad = S.need.get<int>(4) + 3;
                    k_leaf_keys<<<blocks(n0), 256, 0, st>>>(S.node_of.as<int>(), order, n0, rank, S.sb.as<int>(),
                                                            T.level.as<int>(), irr, lkey, bad);
                    int bits = 1;
                    while (bits < 32 && (1u << bits) <= (unsigned)n0) bits++;
                    sort_points_u32(lkey, lkeys, idx, slots, n0, bits, S.tmp, st);
                }
                T.num_slots = n0;
                T.dropped = n - n0;
                T.sorted_build = true;
                copy_host_mirrors(T, st);
                return;
            }
        }
    }
    build_tree_levels(T, S, d_xyz, n, root_c, thr, max_depth, st);
    copy_host_mirrors(T, st);
}

}  // namespace fmm

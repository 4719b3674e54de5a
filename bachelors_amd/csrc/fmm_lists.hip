// Interaction lists on the GPU: the reference's dual-tree recursion
// (FastMultipoleSolver::formInteractionMaps, fast_multipole_solver.cpp:105-165, MAC
// checkIfFarEnough :167-184) evaluated breadth-first over a frontier of (source,
// target) node pairs.  Each pair's fate is a pure function of the two nodes, so the far
// and close sets equal the recursive ones; they are then sorted (and de-duplicated, as
// the reference's std::set does).  The MAC is evaluated in double with the reference's
// operation order and no fused multiply-add, so ties resolve identically.
#include "fmm_lists.hpp"
#include "fmm_prims.hpp"

namespace fmm {

// One 32-byte record per node (centre, first child, level and emptiness) so a node is one
// sector read in the traversal.
struct NodeRec {
    double c[3];
    int fc;     // first child or -1
    int info;   // level | 32 if the node holds points (leaf with npts > 0)
};
struct RadiusTable {
    double r[MAX_LEVELS];
};

__global__ void k_node_records(const double* __restrict__ center, const int* __restrict__ first_child,
                               const int* __restrict__ npts, const int* __restrict__ level, int n,
                               NodeRec* __restrict__ rec) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    NodeRec r;
    r.c[0] = center[3 * i];
    r.c[1] = center[3 * i + 1];
    r.c[2] = center[3 * i + 2];
    r.fc = first_child[i];
    r.info = level[i] | (npts[i] > 0 ? 32 : 0);
    rec[i] = r;
}

__device__ __forceinline__ int lvl(const NodeRec& r) { return r.info & 31; }
__device__ __forceinline__ bool nonempty(const NodeRec& r) { return r.fc >= 0 || (r.info & 32); }
__device__ __forceinline__ bool useful(const NodeRec& r) { return r.fc < 0 && (r.info & 32); }

__device__ __forceinline__ bool far_enough(const NodeRec& s, const RadiusTable& RS, const NodeRec& t,
                                           const RadiusTable& RT, double mac_scale) {
    double r1 = RS.r[lvl(s)], r2 = RT.r[lvl(t)];
    double mn = fmin(r1, r2), mx = fmax(r1, r2);
    double md = __dadd_rn(__dmul_rn(2.0, mx), mn);
    md = __dmul_rn(md, mac_scale);
    double dx = __dsub_rn(s.c[0], t.c[0]);
    double dy = __dsub_rn(s.c[1], t.c[1]);
    double dz = __dsub_rn(s.c[2], t.c[2]);
    double d2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
    return __dmul_rn(md, md) < d2;
}

// fate of one (s, t) pair of the traversal (fast_multipole_solver.cpp:105-165): far, close, or
// expanded into the non-empty children of s (desc_s) or of t (mask: which children)
struct Fate {
    int nfar, nclose, nnext, fc, mask;
    bool desc_s;
};
__device__ __forceinline__ Fate pair_fate(int s, int t, const NodeRec* __restrict__ S, const RadiusTable& RS,
                                          const NodeRec* __restrict__ T, const RadiusTable& RT, double mac_scale) {
    Fate f = {0, 0, 0, -1, 0, false};
    const NodeRec rs = S[s], rt = T[t];
    if (far_enough(rs, RS, rt, RT, mac_scale)) {
        f.nfar = 1;
        return f;
    }
    if (useful(rt)) {
        bool close_enough = true;
        if (rs.fc >= 0)
            for (int o = 0; o < 8; o++)
                if (far_enough(S[rs.fc + o], RS, rt, RT, mac_scale)) { close_enough = false; break; }
        if (close_enough) {
            f.nclose = 1;
            return f;
        }
        f.desc_s = true;
        f.fc = rs.fc;
    } else {
        double radt = RT.r[lvl(rt)], rads = RS.r[lvl(rs)];
        if ((radt < rads || useful(rt)) && !useful(rs)) { f.desc_s = true; f.fc = rs.fc; }
        else f.fc = rt.fc;
    }
    for (int o = 0; o < 8; o++)
        if (f.desc_s ? nonempty(S[f.fc + o]) : nonempty(T[f.fc + o])) f.mask |= 1 << o;
    f.nnext = __popc(f.mask);
    return f;
}

// exclusive scan of three counts over the workgroup (256 threads); returns the totals
__device__ __forceinline__ void block_scan3(unsigned& a, unsigned& b, unsigned& c, unsigned tot[3]) {
    __shared__ unsigned wsum[4][3];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned ia = a, ib = b, ic = c;
    for (int d = 1; d < 64; d <<= 1) {
        unsigned va = __shfl_up(ia, d, 64), vb = __shfl_up(ib, d, 64), vc = __shfl_up(ic, d, 64);
        if (lane >= d) { ia += va; ib += vb; ic += vc; }
    }
    if (lane == 63) { wsum[w][0] = ia; wsum[w][1] = ib; wsum[w][2] = ic; }
    __syncthreads();
    unsigned oa = 0, ob = 0, oc = 0;
    for (int k = 0; k < w; k++) { oa += wsum[k][0]; ob += wsum[k][1]; oc += wsum[k][2]; }
    tot[0] = wsum[0][0] + wsum[1][0] + wsum[2][0] + wsum[3][0];
    tot[1] = wsum[0][1] + wsum[1][1] + wsum[2][1] + wsum[3][1];
    tot[2] = wsum[0][2] + wsum[1][2] + wsum[2][2] + wsum[3][2];
    a = oa + ia - a;
    b = ob + ib - b;
    c = oc + ic - c;
}

// One traversal step over the frontier (s, t) written by the previous step; its size is read
// on the device (*front_cnt, capped at front_cap), so steps are queued without host syncs and
// an exhausted frontier makes the remaining steps no-ops.  Each workgroup takes a contiguous
// chunk: pass 1 counts its far/close/next outputs, one atomic per counter reserves space,
// pass 2 recomputes the fates and writes.  Far/close outputs are (t, s).
__global__ __launch_bounds__(256) void k_lists_step(const int2* __restrict__ front, const unsigned* __restrict__ front_cnt,
                                                    unsigned front_cap, const NodeRec* __restrict__ S, RadiusTable RS,
                                                    const NodeRec* __restrict__ T, RadiusTable RT, double mac_scale,
                                                    int2* far_out, unsigned* far_cnt, unsigned far_cap,
                                                    int2* close_out, unsigned* close_cnt, unsigned close_cap,
                                                    int2* next, unsigned* next_cnt, unsigned next_cap) {
    __shared__ unsigned base[3];
    const unsigned nf = min(*front_cnt, front_cap);
    const unsigned per = (nf + gridDim.x - 1) / gridDim.x;
    const unsigned b0 = blockIdx.x * per, b1 = min(nf, b0 + per);
    if (b0 >= b1) return;   // uniform over the workgroup
    unsigned cf = 0, cc = 0, cn = 0;
    for (unsigned i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
        const Fate f = pair_fate(front[i].x, front[i].y, S, RS, T, RT, mac_scale);
        cf += f.nfar;
        cc += f.nclose;
        cn += f.nnext;
    }
    unsigned tot[3];
    block_scan3(cf, cc, cn, tot);
    if (threadIdx.x == 0) {
        base[0] = tot[0] ? atomicAdd(far_cnt, tot[0]) : 0;
        base[1] = tot[1] ? atomicAdd(close_cnt, tot[1]) : 0;
        base[2] = tot[2] ? atomicAdd(next_cnt, tot[2]) : 0;
    }
    __syncthreads();
    unsigned pf = base[0] + cf, pc = base[1] + cc, pn = base[2] + cn;
    for (unsigned i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
        const int s = front[i].x, t = front[i].y;
        const Fate f = pair_fate(s, t, S, RS, T, RT, mac_scale);
        if (f.nfar) { if (pf < far_cap) far_out[pf] = make_int2(t, s); pf++; }
        if (f.nclose) { if (pc < close_cap) close_out[pc] = make_int2(t, s); pc++; }
        for (int o = 0; o < 8; o++)
            if (f.mask & (1 << o)) {
                if (pn < next_cap) next[pn] = f.desc_s ? make_int2(f.fc + o, t) : make_int2(s, f.fc + o);
                pn++;
            }
    }
}

// Far-pair sort key: class code (level difference and centre offset, each axis in units of
// the finer box's half size along that axis) above the target index (target_bits wide);
// value = source.  hs / ht: per-level half sizes [level][3] of the source / target boxes.
// An offset that is not a whole multiple of the half size (to 1e-9, relative) sets error bit 2
// (such pairs cannot share an operator); offsets or level differences beyond the compact code
// (|n| <= 127, |d| <= 31) set bit 1, and the classes are then keyed by the wide code below.
constexpr unsigned ERR_RANGE = 1u, ERR_LATTICE = 2u;
constexpr int WIDE_BITS = 19;   // per axis in the wide code: |n| < 2^18
__device__ __forceinline__ long long lattice_offset(double d, double h, unsigned* err) {
    const double q = d / h, r = rint(q);
    if (!(fabs(q - r) <= 1e-9 * fmax(1.0, fabs(q))) || !(fabs(r) < (double)(1 << (WIDE_BITS - 1)))) {
        atomicOr(err, ERR_LATTICE);
        return 0;
    }
    if (r < -127.0 || r > 127.0) atomicOr(err, ERR_RANGE);
    return (long long)r;
}
// the wide class code: (d + 32) above three WIDE_BITS-bit offsets (63 bits)
__device__ __forceinline__ uint64_t wide_code(int d, long long nx, long long ny, long long nz) {
    const long long o = 1ll << (WIDE_BITS - 1);
    return ((uint64_t)(d + 32) << (3 * WIDE_BITS)) | ((uint64_t)(nx + o) << (2 * WIDE_BITS)) |
           ((uint64_t)(ny + o) << WIDE_BITS) | (uint64_t)(nz + o);
}
__global__ void k_far_keys(const int2* __restrict__ far, int n, const NodeRec* __restrict__ S,
                           const NodeRec* __restrict__ T, const double* __restrict__ hs, const double* __restrict__ ht,
                           int use_classes, int target_bits, uint64_t* __restrict__ key, int* __restrict__ val,
                           unsigned* __restrict__ err) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int t = far[i].x, s = far[i].y;
    uint64_t code = 0;
    if (use_classes) {
        const NodeRec rs = S[s], rt = T[t];
        int ls = lvl(rs), lt = lvl(rt);
        // finer of the two boxes (physical size, not tree level), per axis
        long long nx = lattice_offset(rs.c[0] - rt.c[0], fmin(hs[3 * ls], ht[3 * lt]), err);
        long long ny = lattice_offset(rs.c[1] - rt.c[1], fmin(hs[3 * ls + 1], ht[3 * lt + 1]), err);
        long long nz = lattice_offset(rs.c[2] - rt.c[2], fmin(hs[3 * ls + 2], ht[3 * lt + 2]), err);
        int d = ls - lt;
        if (d < -31 || d > 31) atomicOr(err, ERR_RANGE);
        code = ((uint64_t)(d + 32) << 24) | ((uint64_t)((nx + 128) & 255) << 16) | ((uint64_t)((ny + 128) & 255) << 8) |
               (uint64_t)((nz + 128) & 255);
    }
    key[i] = (code << target_bits) | (uint32_t)t;
    val[i] = s;
}

// the wide path: per far pair (unsorted, (t, s)) its wide class code and its target as a key
__global__ void k_far_keys_wide(const int2* __restrict__ far, int n, const NodeRec* __restrict__ S,
                                const NodeRec* __restrict__ T, const double* __restrict__ hs,
                                const double* __restrict__ ht, uint64_t* __restrict__ code, uint32_t* __restrict__ tkey,
                                int* __restrict__ idx, unsigned* __restrict__ err) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = far[i].x, s = far[i].y;
    const NodeRec rs = S[s], rt = T[t];
    const int ls = lvl(rs), lt = lvl(rt);
    const long long nx = lattice_offset(rs.c[0] - rt.c[0], fmin(hs[3 * ls], ht[3 * lt]), err);
    const long long ny = lattice_offset(rs.c[1] - rt.c[1], fmin(hs[3 * ls + 1], ht[3 * lt + 1]), err);
    const long long nz = lattice_offset(rs.c[2] - rt.c[2], fmin(hs[3 * ls + 2], ht[3 * lt + 2]), err);
    code[i] = wide_code(ls - lt, nx, ny, nz);
    tkey[i] = (uint32_t)t;
    idx[i] = i;
}
__global__ void k_gather_u64(const uint64_t* __restrict__ in, const int* __restrict__ idx, int n,
                             uint64_t* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[idx[i]];
}
__global__ void k_gather_pairs(const int2* __restrict__ far, const int* __restrict__ idx, int n, int* __restrict__ ft,
                               int* __restrict__ fs) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int2 p = far[idx[i]];
    ft[i] = p.x;
    fs[i] = p.y;
}
__global__ void k_seg_flags_u64(const uint64_t* __restrict__ code, int n, uint32_t* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || code[i] != code[i - 1]) ? 1u : 0u;
}

// close pairs keyed (target, source) with the source in the low sbits bits
__global__ void k_close_keys(const int2* __restrict__ close, int n, int sbits, uint64_t* __restrict__ key,
                             int* __restrict__ val) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = ((uint64_t)(uint32_t)close[i].x << sbits) | (uint32_t)close[i].y;
    val[i] = close[i].y;
}


// sorted keys back to (target, source[, class code]): far keys carry the target in the low
// target_bits (class code above), close keys in the high word
__global__ void k_unpack_pairs(const uint64_t* __restrict__ key, const int* __restrict__ val, int n, int target_shift,
                               int target_bits, int* __restrict__ out_t, int* __restrict__ out_s,
                               uint32_t* __restrict__ out_code) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out_t[i] = (int)(uint32_t)((key[i] >> target_shift) & ((1ull << target_bits) - 1));
    out_s[i] = val[i];
    if (out_code) out_code[i] = (uint32_t)(key[i] >> target_bits);
}

// the class count (scan total) and the key-range error flag in one readback
__global__ void k_class_count(const uint32_t* __restrict__ pos, const uint32_t* __restrict__ flag, int n,
                              const unsigned* __restrict__ err, unsigned* __restrict__ out) {
    out[0] = pos[n - 1] + flag[n - 1];
    out[1] = *err;
}

// segment boundaries: class starts (far) / target starts (close)
__global__ void k_seg_flags(const uint32_t* __restrict__ code, int n, uint32_t* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || code[i] != code[i - 1]) ? 1u : 0u;
}
__global__ void k_seg_starts(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, int n,
                             int* __restrict__ starts) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    starts[pos[i]] = i;
}
__global__ void k_target_ranges(const int* __restrict__ ct, int n, int* __restrict__ cbeg, int* __restrict__ cend) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int t = ct[i];
    if (i == 0 || ct[i - 1] != t) cbeg[t] = i;
    if (i == n - 1 || ct[i + 1] != t) cend[t] = i + 1;
}
__global__ void k_fill_int(int* p, int n, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}


static RadiusTable radii(const TreeDev& T) {
    RadiusTable r;
    for (int k = 0; k < MAX_LEVELS; k++) r.r[k] = T.radius[k];
    return r;
}

static uint32_t read_u32(const void* p, hipStream_t st) {
    uint32_t v;
    FMM_HIP(hipMemcpyAsync(&v, p, 4, hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
    return v;
}

static int bits_for(int n) {   // smallest b with 2^b >= n (at least 1)
    int b = 1;
    while (b < 31 && (1 << b) < n) b++;
    return b;
}

// sort by the low end_bit bits of the 64-bit key and unpack.  The traversal reaches every
// (source, target) pair at most once (each pair has a unique parent pair), so the sets need no
// de-duplication, unlike the reference's std::set inserts.
static void sort_unpack(ListScratch& L, int n, int end_bit, int target_shift, int target_bits, int* out_t, int* out_s,
                        uint32_t* out_code, hipStream_t st) {
    if (n == 0) return;
    uint64_t* k2 = L.key2.get<uint64_t>(n);
    int* v2 = L.val2.get<int>(n);
    sort_pairs_u64(L.key.as<uint64_t>(), k2, L.val.as<int>(), v2, n, end_bit, L.tmp, st);
    k_unpack_pairs<<<blocks(n), 256, 0, st>>>(k2, v2, n, target_shift, target_bits, out_t, out_s, out_code);
}

void build_lists(Lists& out, ListScratch& L, const TreeDev& S, const TreeDev& T, double mac_scale, bool classes,
                 hipStream_t st, hipEvent_t far_ready, const std::function<void()>& after_far) {
    RadiusTable rs = radii(S), rt = radii(T);
    NodeRec* srec = L.srec.get<NodeRec>(std::max(S.num_nodes, 1));
    NodeRec* trec = L.trec.get<NodeRec>(std::max(T.num_nodes, 1));
    if (S.num_nodes)
        k_node_records<<<blocks(S.num_nodes), 256, 0, st>>>(S.center.as<double>(), S.first_child.as<int>(),
                                                            S.npts.as<int>(), S.level.as<int>(), S.num_nodes, srec);
    if (T.num_nodes)
        k_node_records<<<blocks(T.num_nodes), 256, 0, st>>>(T.center.as<double>(), T.first_child.as<int>(),
                                                            T.npts.as<int>(), T.level.as<int>(), T.num_nodes, trec);
    // every step descends one of the two trees by a level: the traversal ends within this many
    const int max_steps = std::max(S.num_levels, 1) + std::max(T.num_levels, 1);
    // counters: far, close, error flag, then one frontier count per step
    unsigned* cnt = L.counters.get<unsigned>(4 + max_steps + 1 + 2);
    unsigned far_cap = std::max<unsigned>(L.far_cap, 1u << 16), close_cap = std::max<unsigned>(L.close_cap, 1u << 16),
             front_cap = std::max<unsigned>(L.front_cap, 1u << 16);
    for (int attempt = 0; attempt < 8; attempt++) {
        int2* far = L.far.get<int2>(far_cap);
        int2* close = L.close.get<int2>(close_cap);
        int2* fa = L.frontA.get<int2>(front_cap);
        int2* fb = L.frontB.get<int2>(front_cap);
        FMM_HIP(hipMemsetAsync(cnt, 0, (4 + max_steps + 1) * sizeof(unsigned), st));
        // the reference returns at once if either root is empty (fast_multipole_solver.cpp:206-210)
        bool s_ok = !S.h_first_child.empty() && (S.h_first_child[0] >= 0 || S.h_npts[0] > 0);
        bool t_ok = !T.h_first_child.empty() && (T.h_first_child[0] >= 0 || T.h_npts[0] > 0);
        if (s_ok && t_ok) {
            L.h_seed[0] = 0;
            L.h_seed[1] = 0;
            L.h_seed[2] = 1;
            FMM_HIP(hipMemcpyAsync(fa, L.h_seed, sizeof(int2), hipMemcpyHostToDevice, st));
            FMM_HIP(hipMemcpyAsync(cnt + 4, L.h_seed + 2, sizeof(unsigned), hipMemcpyHostToDevice, st));
            // steps queue without host synchronisation; the grid strides over any frontier size.
            // Step k's frontier holds at most 8^k pairs (each pair descends into at most 8), so the
            // first steps launch only the workgroups they can use (they start sooner on a busy GPU)
            size_t bound = 1;
            for (int k = 0; k < max_steps; k++) {
                const int grid = (int)std::min<size_t>(1024, (bound + 255) / 256);
                bound = std::min<size_t>(bound * 8, (size_t)front_cap);
                k_lists_step<<<grid, 256, 0, st>>>(fa, cnt + 4 + k, front_cap, srec, rs, trec, rt, mac_scale, far,
                                                   cnt + 0, far_cap, close, cnt + 1, close_cap, fb, cnt + 4 + k + 1,
                                                   front_cap);
                std::swap(fa, fb);
            }
        }
        std::vector<unsigned> h(4 + max_steps + 1);
        FMM_HIP(hipMemcpyAsync(h.data(), cnt, h.size() * sizeof(unsigned), hipMemcpyDeviceToHost, st));
        FMM_HIP(hipStreamSynchronize(st));
        unsigned fmax = 0;
        int steps = 0;
        for (int k = 0; k <= max_steps; k++) {
            fmax = std::max(fmax, h[4 + k]);
            if (h[4 + k]) steps = k + 1;
        }
        if (h[4 + max_steps] != 0) throw Error(4, "interaction-list traversal did not terminate");
        if (h[0] > far_cap || h[1] > close_cap || fmax > front_cap) {
            far_cap = std::max(far_cap, 2 * h[0]);
            close_cap = std::max(close_cap, 2 * h[1]);
            front_cap = std::max(front_cap, 2 * fmax);
            continue;
        }
        L.far_cap = far_cap;
        L.close_cap = close_cap;
        L.front_cap = front_cap;
        int nfar = (int)h[0], nclose = (int)h[1];
        out.bfs_steps = steps;
        // ---- far pairs: sort by (class, target), de-duplicate, segment into classes
        {
            uint64_t* key = L.key.get<uint64_t>(std::max(nfar, 1));
            int* val = L.val.get<int>(std::max(nfar, 1));
            unsigned* err = cnt + 2;
            double* dh = L.hs.get<double>(6 * MAX_LEVELS);
            for (int k = 0; k < MAX_LEVELS; k++)
                for (int a = 0; a < 3; a++) {
                    L.h_half[3 * k + a] = S.half[k][a];
                    L.h_half[3 * (MAX_LEVELS + k) + a] = T.half[k][a];
                }
            FMM_HIP(hipMemcpyAsync(dh, L.h_half, 6 * MAX_LEVELS * sizeof(double), hipMemcpyHostToDevice, st));
            const int tbits = bits_for(T.num_nodes);
            if (nfar > 0)
                k_far_keys<<<blocks(nfar), 256, 0, st>>>(far, nfar, srec, trec, dh, dh + 3 * MAX_LEVELS, classes ? 1 : 0,
                                                         tbits, key, val, err);
            int* ft = out.far_t.get<int>(std::max(nfar, 1));
            int* fs = out.far_s.get<int>(std::max(nfar, 1));
            uint32_t* code = out.far_code.get<uint32_t>(std::max(nfar, 1));
            sort_unpack(L, nfar, tbits + (classes ? 30 : 0), 0, tbits, ft, fs, code, st);
            out.num_far = nfar;
            // classes
            out.num_classes = 0;
            out.h_class_begin.clear();
            if (nfar > 0 && classes) {
                uint32_t* flag = L.flag.get<uint32_t>(nfar);
                uint32_t* pos = L.pos.get<uint32_t>(nfar);
                k_seg_flags<<<blocks(nfar), 256, 0, st>>>(code, nfar, flag);
                exclusive_scan_u32(flag, pos, nfar, L.tmp, st);
                unsigned* cc = cnt + 4 + max_steps + 1;
                k_class_count<<<1, 1, 0, st>>>(pos, flag, nfar, err, cc);
                unsigned hc[2];
                FMM_HIP(hipMemcpyAsync(hc, cc, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, st));
                FMM_HIP(hipStreamSynchronize(st));
                if (hc[1] & ERR_LATTICE)
                    throw Error(4, "interaction offset off the box lattice (the shared frame needs lattice-aligned boxes)");
                if (hc[1] & ERR_RANGE) {
                    // offsets or level differences beyond the compact code (deep, strongly adaptive
                    // trees): classes by the wide code, sorted in two stable passes (by target, then
                    // by code), so the order is still (class, target)
                    int* perm = L.perm.get<int>(nfar);
                    int* perm2 = L.perm2.get<int>(nfar);
                    uint64_t* wc = L.wkey.get<uint64_t>(nfar);
                    uint64_t* wc2 = L.wkey2.get<uint64_t>(nfar);
                    uint32_t* tk = L.tkey.get<uint32_t>(nfar);
                    uint32_t* tk2 = L.tkey2.get<uint32_t>(nfar);
                    int* idx = L.idx.get<int>(nfar);
                    FMM_HIP(hipMemsetAsync(err, 0, sizeof(unsigned), st));
                    k_far_keys_wide<<<blocks(nfar), 256, 0, st>>>(far, nfar, srec, trec, dh, dh + 3 * MAX_LEVELS, wc, tk,
                                                                  idx, err);
                    sort_pairs_u32(tk, tk2, idx, perm, nfar, tbits, L.tmp, st);
                    k_gather_u64<<<blocks(nfar), 256, 0, st>>>(wc, perm, nfar, wc2);
                    sort_pairs_u64(wc2, wc, perm, perm2, nfar, 3 * WIDE_BITS + 6, L.tmp, st);
                    k_gather_pairs<<<blocks(nfar), 256, 0, st>>>(far, perm2, nfar, ft, fs);
                    k_seg_flags_u64<<<blocks(nfar), 256, 0, st>>>(wc, nfar, flag);
                    exclusive_scan_u32(flag, pos, nfar, L.tmp, st);
                    k_class_count<<<1, 1, 0, st>>>(pos, flag, nfar, err, cc);
                    FMM_HIP(hipMemcpyAsync(hc, cc, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, st));
                    FMM_HIP(hipStreamSynchronize(st));
                    if (hc[1] & ERR_LATTICE) throw Error(4, "interaction offset beyond the wide class code");
                }
                int nc = (int)hc[0];
                int* starts = out.class_begin.get<int>(nc + 1);
                k_seg_starts<<<blocks(nfar), 256, 0, st>>>(flag, pos, nfar, starts);
                out.h_class_begin.resize(nc + 1);
                FMM_HIP(hipMemcpyAsync(out.h_class_begin.data(), starts, nc * sizeof(int), hipMemcpyDeviceToHost, st));
                FMM_HIP(hipStreamSynchronize(st));
                out.h_class_begin[nc] = nfar;
                out.num_classes = nc;
            }
        }
        if (far_ready) FMM_HIP(hipEventRecord(far_ready, st));
        if (after_far) after_far();
        // ---- close pairs: sort by (target, source), per-target ranges
        {
            uint64_t* key = L.key.get<uint64_t>(std::max(nclose, 1));
            int* val = L.val.get<int>(std::max(nclose, 1));
            // (target, source) in tbits + sbits bits: the radix sort runs over those bits only
            const int tb = bits_for(T.num_nodes), sb = bits_for(S.num_nodes);
            if (nclose > 0) k_close_keys<<<blocks(nclose), 256, 0, st>>>(close, nclose, sb, key, val);
            int* ct = out.close_t.get<int>(std::max(nclose, 1));
            int* cs = out.close_s.get<int>(std::max(nclose, 1));
            sort_unpack(L, nclose, tb + sb, sb, tb, ct, cs, nullptr, st);
            out.num_close = nclose;
            int* cb = out.close_begin.get<int>(std::max(T.num_nodes, 1));
            int* ce = out.close_end.get<int>(std::max(T.num_nodes, 1));
            k_fill_int<<<blocks(T.num_nodes), 256, 0, st>>>(cb, T.num_nodes, 0);
            k_fill_int<<<blocks(T.num_nodes), 256, 0, st>>>(ce, T.num_nodes, 0);
            if (nclose > 0) k_target_ranges<<<blocks(nclose), 256, 0, st>>>(ct, nclose, cb, ce);
        }
        return;
    }
    throw Error(4, "interaction list buffers did not converge");
}

}  // namespace fmm

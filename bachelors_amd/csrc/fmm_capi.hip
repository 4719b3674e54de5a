// C ABI and stage orchestration of the MI355X Biot-Savart FMM (include/fmm_capi.h).
//
// Stage order follows FastMultipoleSolver (fast_multipole_solver.cpp): trees, leaf
// multipoles (P2M), upward pass (M2M), downward pass (interaction lists, M2L, L2L),
// evaluation (L2P curl + near field).  Everything after the inputs is device resident:
// the host only plans per-level pair lists and GEMM tiles from small node tables.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <exception>
#include <mutex>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fmm_capi.h"
#include "fmm_adjoint.hpp"
#include "fmm_expand.hpp"
#include "fmm_gemm.hpp"
#include "fmm_lists.hpp"
#include "fmm_prims.hpp"
#include "fmm_tables.hpp"
#include "fmm_tree.hpp"
#include "fmm_treecode.hpp"

using namespace fmm;

namespace {

// Device scratch of the allocation-free direct sum (fmm_direct_B_noalloc_dev): the partial sums
// of its source ranges, FMM_NOALLOC_SCRATCH_DOUBLES doubles (32 MiB) per GPU, held until the
// process ends.  Host code never passes this name to a kernel: the device address is looked up
// once per device (scratch_on_current_device) and that pointer is used.
__device__ __align__(256) double scratch[FMM_NOALLOC_SCRATCH_DOUBLES];

std::mutex g_scratch_mutex;
std::vector<double*> g_scratch_ptr;   // per device id

// the device address of `scratch` on the current device (looked up on the first call per device)
double* scratch_on_current_device() {
    int dev = 0;
    FMM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        FMM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw Error(FMM_ERR_HIP, "no device address for the scratch array");
        g_scratch_ptr[dev] = (double*)ptr;
    }
    return g_scratch_ptr[dev];
}

enum Stage { ST_EMPTY = 0, ST_TREES = 1, ST_LEAVES = 2, ST_UP = 3, ST_DOWN = 4, ST_EVAL = 5 };
// M2M levels with at least this many parent columns use parent-major tiles
constexpr int M2M_PARENT_MAJOR_COLUMNS = 64 * 256;

// stage events: the stream-A chain, plus start marks of the stages that may run on the side
// stream (lists, P2P) and of the translations after the lists (E_DOWN0)
enum Ev { E_START, E_TREE, E_P2M, E_M2M, E_LISTS, E_TABLES, E_M2L, E_L2L, E_L2P, E_P2P, E_END, E_M2L_GEMM,
          E_LISTS0, E_DOWN0, E_P2P0, E_GEMM0, E_FAR, E_COUNT };

// Launch data built on the host for each evaluation (pair lists, tiles, operator specs) is
// appended to a pinned buffer and sent with one copy per stage: per-level copies from pageable
// memory each cost a blit launch on the stream.  begin() waits for the previous copy out of
// the host buffer before it is overwritten.
struct Staging {
    char* h = nullptr;
    size_t cap = 0, used = 0;
    DevBuf d;
    hipEvent_t copied = nullptr;
    Staging() = default;
    Staging(const Staging&) = delete;
    Staging& operator=(const Staging&) = delete;
    ~Staging() {
        if (h) (void)hipHostFree(h);
        if (copied) (void)hipEventDestroy(copied);
    }
    void begin() {
        if (copied) FMM_HIP(hipEventSynchronize(copied));
        used = 0;
    }
    // byte offset of a copy of bytes at p (16-byte aligned)
    size_t put(const void* p, size_t bytes) {
        size_t off = (used + 15) & ~(size_t)15;
        if (off + bytes > cap) {
            size_t nc = std::max(off + bytes, 2 * cap + (1 << 16));
            char* nh = nullptr;
            FMM_HIP(hipHostMalloc((void**)&nh, nc, hipHostMallocDefault));
            if (used) std::memcpy(nh, h, used);
            if (h) FMM_HIP(hipHostFree(h));
            h = nh;
            cap = nc;
        }
        if (bytes) std::memcpy(h + off, p, bytes);
        used = off + bytes;
        return off;
    }
    template <class T> size_t put(const std::vector<T>& v) { return put(v.data(), v.size() * sizeof(T)); }
    // one copy of everything put since begin(); returns the device base of the offsets
    char* send(hipStream_t st) {
        char* dev = (char*)d.reserve(std::max<size_t>(used, 16));
        if (used) FMM_HIP(hipMemcpyAsync(dev, h, used, hipMemcpyHostToDevice, st));
        if (!copied) FMM_HIP(hipEventCreateWithFlags(&copied, hipEventDisableTiming));
        FMM_HIP(hipEventRecord(copied, st));
        return dev;
    }
};

struct Pass {               // one batched GEMM launch: tiles over (src, dst) pairs (staged data)
    int ntiles = 0, npairs = 0;
    int width = 64;         // columns per tile
    size_t o_src = 0, o_dst = 0, o_tiles = 0;
    const int* src = nullptr;
    const int* dst = nullptr;
    const GemmTile* tiles = nullptr;
};

struct OpTables {           // FP64 master + the precision actually applied
    DevBuf f64, f32, hi, lo, specs;
    int nops = 0;
};

}  // namespace

struct fmm_ctx {
    fmm_config cfg;
    std::string err;
    double* scratch = nullptr;                   // this device's address of the no-alloc scratch array
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int p = 10, L = 121, Lp = 128;
    Precision prec = Precision::F64;
    // inputs (input order)
    int ns = 0, nt = 0;
    bool have_J = false;
    bool bem = false;                            // Problem::BEM sources (B and normal per point)
    DevBuf src_B, src_N, sbd;                    // BEM inputs (input order); w (B . n) per slot
    double tmin[3] = {1e+16, 1e+16, 1e+16}, tmax[3] = {-1e+16, -1e+16, -1e+16};   // target extent
    DevBuf src_xyz, src_w, src_J, tgt_xyz;
    // trees
    TreeDev S, T;
    TreeScratch tscr, tscr_t;                    // source / target tree scratch (built concurrently)
    ScaleTable sigS, sigT;
    DevBuf sxyz, sq, sJ, txyz;                   // slot order
    DevBuf nP0, nP1, nP2;                        // near-field sources packed per leaf (slot order)
    PinnedVec<int> h_src_leaves, h_tgt_leaves;   // useful (non-empty) leaves
    DevBuf src_leaves, tgt_leaves;
    // expansions [node][3][Lp]
    DevBuf M, Lx;
    // translations
    OpTables up_ops, down_ops;
    std::vector<Pass> m2m_pass, l2l_pass, m2m_adj_pass;   // m2m_adj_pass: per orientation, for M2M^T
    Pass m2l_pass;
    // their launch data; operator specs go first in a copy of their own, so the table kernels
    // run while the host plans the passes
    Staging up_specs, up_stage, down_stage, m2l_specs, m2l_stage, adj_stage;
    DevBuf pp_tiles;                             // per-pair path: tiles of one chunk
    Lists lists;
    ListScratch lscr;
    PairIndex m2l_ix;                            // far pairs by target node
    // F16X3: per-coefficient scales dm, the M2L tables split in FP16 (their FP32 staging), the
    // multipole column scales; down_l2l = L2L operators ahead of the M2L ones in the down tables
    DevBuf dm, m2l_hi, m2l_lo, m2l_rinv, m2l_work, cscale, Mhl;   // Mhl: split multipoles (hi | lo)
    DevBuf queues;   // work counters: [0, 8) the M2L GEMM's per-XCD tiles, [8] the near-field leaves
    int down_l2l = 0;
    const OpSpec* down_specs = nullptr;          // device specs of the down tables
    DevBuf m2l_slab;                             // M2L result per (pair, component)
    // evaluation
    DevBuf Bfar, Bnear, Bout;
    DevBuf Afar, Anear, Aout;                    // vector potential (calcA)
    DevBuf direct_part, pts, tc_B, tc_A;         // direct-sum partials; host query points and results
    double current = 1.0;
    // adjoint state
    AdjointState adj;
    int stage = ST_EMPTY;
    hipEvent_t ev[E_COUNT];
    bool ev_ok = false;
    // side stream of the fused evaluation (lists and P2P beside the multipole chain) and its joins
    hipStream_t side = nullptr;
    hipEvent_t j_lists = nullptr, j_pack = nullptr, j_p2p = nullptr, j_gemm = nullptr,
               j_start = nullptr, j_tables = nullptr,   // j_tables: everything the side stream builds for M2L
               j_src = nullptr, j_tgt = nullptr,        // source tree on the main / target side on the side stream
               j_caller = nullptr;                      // the caller's default-stream work before a _dev read
    long long near_inter = -1;
    bool lists_ready = false;
    bool leaves_split = false;   // F16X3: the last P2M also wrote its leaves' split M2L columns

    bool f64() const { return prec == Precision::F64; }
    // precision of the M2M / L2L passes: F16X3 splits only the M2L, the others run in FP32
    Precision pass_prec() const { return prec == Precision::F16X3 ? Precision::F32 : prec; }
    size_t esize() const { return f64() ? sizeof(double) : sizeof(float); }
    void rec(int e) { if (ev_ok) FMM_HIP(hipEventRecord(ev[e], stream)); }
    void rec_on(int e, hipStream_t s) { if (ev_ok) FMM_HIP(hipEventRecord(ev[e], s)); }
};

namespace {

template <class F> int guard(fmm_ctx* ctx, F&& f) {
    if (!ctx) return FMM_ERR_ARG;
    try {
        f();
        ctx->err.clear();
        return FMM_OK;
    } catch (const Error& e) {
        ctx->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return FMM_ERR_HIP;
    }
}

inline int blocksN(long long n) { return (int)((n + 255) / 256); }

void require(bool c, int code, const char* msg) {
    if (!c) throw Error(code, msg);
}

// Device-pointer entry points read caller memory on the context stream.  The caller's work on
// the default (null) stream is ordered before it by an event (no host wait, no device-wide
// synchronisation); work on other caller streams is the caller's to order (fmm_set_stream).
void join_default_stream(fmm_ctx* c) {
    FMM_HIP(hipEventRecord(c->j_caller, nullptr));
    FMM_HIP(hipStreamWaitEvent(c->stream, c->j_caller, 0));
}

bool nonempty_host(const TreeDev& T, int n) { return T.h_first_child[n] >= 0 || T.h_npts[n] > 0; }

// columns per M2L tile: the FP16 split kernel takes 128 (two 16-column blocks per wave, each
// operator fragment read from LDS once for both; measured 2.37 -> 2.09 ms on the 1M workload)
int m2l_tile_cols(const fmm_ctx* c) { return c->prec == Precision::F16X3 ? 128 : 64; }

// F16X3: the scale of real coefficient i = (l, m), a power of 2 with dm_i max |R_i(u)| in
// (0.5, 1] over |u| <= sqrt(3) (the scaled position of any point in its box), so every scaled
// multipole coefficient is bounded by the same sum of |charges|.  The maximum is taken over a
// Fibonacci lattice of 4096 directions at |u| = sqrt(3) (R_i is homogeneous of degree l);
// coefficients beyond L get 1.
std::vector<float> coefficient_scales(int p, int kpad) {
    const int L = (p + 1) * (p + 1);
    std::vector<double> mx(L, 0.0);
    const int N = 4096;
    const double ga = M_PI * (3.0 - std::sqrt(5.0)), rr = std::sqrt(3.0);
    std::vector<double> pr((p + 1) * (p + 1)), pi((p + 1) * (p + 1));
    for (int n = 0; n < N; n++) {
        const double z0 = 1.0 - (2.0 * n + 1.0) / N, s0 = std::sqrt(std::max(0.0, 1.0 - z0 * z0));
        const double x = rr * s0 * std::cos(ga * n), y = rr * s0 * std::sin(ga * n), z = rr * z0, r2 = 3.0;
        for (int m = 0; m <= p; m++) {
            double dr = 1.0, di = 0.0;   // Phi_m^m = (x + iy)^m / (2^m m!)
            for (int i = 1; i <= m; i++) {
                const double nr = (dr * x - di * y) / (2.0 * i), ni = (dr * y + di * x) / (2.0 * i);
                dr = nr;
                di = ni;
            }
            double p1r = dr, p1i = di, p2r = 0.0, p2i = 0.0;
            for (int l = m; l <= p; l++) {
                double vr = dr, vi = di;
                if (l > m) {
                    vr = ((2.0 * l - 1) * z * p1r - r2 * p2r) / (double)(l * l - m * m);
                    vi = ((2.0 * l - 1) * z * p1i - r2 * p2i) / (double)(l * l - m * m);
                    p2r = p1r; p2i = p1i; p1r = vr; p1i = vi;
                }
                const double a = m == 0 ? std::fabs(vr) : std::sqrt(2.0) * std::hypot(vr, vi);
                const int ip = l * l + l + m, im = l * l + l - m;
                mx[ip] = std::max(mx[ip], a);
                mx[im] = std::max(mx[im], a);
            }
        }
    }
    std::vector<float> dm(kpad, 1.0f);
    for (int i = 0; i < L; i++)
        if (mx[i] > 0) dm[i] = (float)std::ldexp(1.0, -(int)std::ceil(std::log2(mx[i])));
    return dm;
}

// reference child order (quadrature_octree.cpp:88-96)
const int SIGN[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                        {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

void stage_pass(Staging& sg, Pass& P, const std::vector<int>& src, const std::vector<int>& dst,
                const std::vector<GemmTile>& tiles) {
    P.npairs = (int)src.size();
    P.ntiles = (int)tiles.size();
    P.o_src = sg.put(src);
    P.o_dst = sg.put(dst);
    P.o_tiles = sg.put(tiles);
    P.src = P.dst = nullptr;
    P.tiles = nullptr;
}
void resolve_pass(Pass& P, const char* dev) {
    P.src = (const int*)(dev + P.o_src);
    P.dst = (const int*)(dev + P.o_dst);
    P.tiles = (const GemmTile*)(dev + P.o_tiles);
}

// O.nops operators from device specs, written in the precision the GEMMs apply
// O.nops operators (from device specs) in the precision of the passes that apply them
void build_tables(fmm_ctx* c, OpTables& O, const OpSpec* d_specs, hipStream_t st, int nops = -1) {
    if (nops < 0) nops = O.nops;
    size_t n = (size_t)nops * c->Lp * c->Lp;
    if (!n) return;
    const Precision pr = c->pass_prec();
    if (pr == Precision::F64)
        make_operators(d_specs, nops, c->p, c->Lp, 0, O.f64.get<double>(n), nullptr, st);
    else if (pr == Precision::F32)
        make_operators(d_specs, nops, c->p, c->Lp, 1, O.f32.get<float>(n), nullptr, st);
    else
        make_operators(d_specs, nops, c->p, c->Lp, 2, O.hi.get<__bf16>(n), O.lo.get<__bf16>(n), st);
}

const void* ops_ptr(const fmm_ctx* c, const OpTables& O) {
    const Precision pr = c->pass_prec();
    if (pr == Precision::F64) return O.f64.ptr;
    if (pr == Precision::F32) return O.f32.ptr;
    return O.hi.ptr;
}

void run_pass(fmm_ctx* c, const OpTables& O, const Pass& P, const void* src_exp, void* dst_exp, int epilogue,
              int tri) {
    if (!P.ntiles) return;
    class_gemm(c->pass_prec(), ops_ptr(c, O), O.lo.ptr, P.tiles, P.ntiles, P.src, P.dst, src_exp, dst_exp, c->Lp,
               epilogue, c->stream, false, 1, tri);
}

// Roots of both trees (tc, th: target root) and the start mark of an evaluation.
void trees_begin(fmm_ctx* c, double tc[3], double th[3]) {
    require(c->ns > 0 || c->nt > 0, FMM_ERR_STATE, "sources and targets are not set");
    hipStream_t st = c->stream;
    c->rec(E_START);
    const fmm_config& g = c->cfg;
    if (g.frame == FMM_FRAME_SHARED) {
        // the source root, grown by doublings along the source lattice until it holds every
        // target (half-open), so all boxes of both trees stay on one dyadic lattice
        const double* c0 = g.source_root_center;
        const double* h0 = g.source_root_half;
        int k = 0;
        for (;; k++) {
            require(k <= 12, FMM_ERR_LIMIT, "targets too far from the source root for the shared frame");
            double sc = std::ldexp(1.0, k);
            bool ok = true;
            for (int a = 0; a < 3 && ok; a++) {
                double cell = 2 * h0[a];
                double j = c->nt ? std::ceil((c0[a] - h0[a] - c->tmin[a]) / cell) : 0.0;
                if (j < 0) j = 0;
                if (j > sc - 1) { ok = false; break; }
                double lo = c0[a] - h0[a] - cell * j;
                if (c->nt && !(c->tmax[a] < lo + cell * sc)) { ok = false; break; }
                th[a] = h0[a] * sc;
                tc[a] = lo + th[a];
            }
            if (ok) break;
        }
    } else {
        // math::getBoundingBox (math.cpp:39-57)
        for (int a = 0; a < 3; a++) {
            double hd = (c->tmax[a] - c->tmin[a]) / 2 + 1e-5;
            tc[a] = c->tmin[a] + hd;
            th[a] = hd * 1.1;
        }
    }
    c->near_inter = -1;
    c->lists_ready = false;
    // the target tree is built on the side stream, which first waits for everything queued on
    // the main stream (a previous evaluation may still read the target tree there)
    FMM_HIP(hipEventRecord(c->j_start, st));
    FMM_HIP(hipStreamWaitEvent(c->side, c->j_start, 0));
}

// useful (non-empty) leaves of T into h (pinned) and d, copied on stream st
void upload_leaves(const TreeDev& T, PinnedVec<int>& h, DevBuf& d, hipStream_t st) {
    h.resize(std::max(T.num_nodes, 1));
    size_t k = 0;
    for (int n = 0; n < T.num_nodes; n++)
        if (T.h_first_child[n] < 0 && T.h_npts[n] > 0) h[k++] = n;
    h.resize(k);
    if (k) FMM_HIP(hipMemcpyAsync(d.get<int>(k), h.data(), k * sizeof(int), hipMemcpyHostToDevice, st));
}

// source tree, sources in slot (leaf DFS) order and the useful source leaves, on stream A
void build_source_side(fmm_ctx* c) {
    hipStream_t st = c->stream;
    const fmm_config& g = c->cfg;
    build_tree(c->S, c->tscr, c->src_xyz.as<double>(), c->ns, g.source_root_center, g.source_root_half,
               g.source_leaf_capacity, g.max_depth, true, st);
    for (int k = 0; k < MAX_LEVELS; k++) c->sigS.s[k] = c->S.scale[k];
    int nss = c->S.num_slots;
    if (c->bem)
        launch_bem_charges(c->S.slots.as<int>(), nss, c->src_xyz.as<double>(), c->src_w.as<double>(),
                           c->src_B.as<double>(), c->src_N.as<double>(),
                           c->sxyz.get<double>(3 * (size_t)std::max(nss, 1)),
                           c->sq.get<double>(3 * (size_t)std::max(nss, 1)), c->sbd.get<double>((size_t)std::max(nss, 1)),
                           st);
    else
        launch_charges(c->S.slots.as<int>(), nss, c->src_xyz.as<double>(), c->src_w.as<double>(),
                       c->have_J ? c->src_J.as<double>() : nullptr, c->sxyz.get<double>(3 * (size_t)std::max(nss, 1)),
                       c->sq.get<double>(3 * (size_t)std::max(nss, 1)), c->sJ.get<double>(3 * (size_t)std::max(nss, 1)),
                       st);
    upload_leaves(c->S, c->h_src_leaves, c->src_leaves, st);
}

// target tree, targets in slot order and the useful target leaves, on stream ts
void build_target_side(fmm_ctx* c, const double tc[3], const double th[3], hipStream_t ts) {
    const fmm_config& g = c->cfg;
    build_tree(c->T, c->tscr_t, c->tgt_xyz.as<double>(), c->nt, tc, th, g.target_leaf_capacity + 1, g.max_depth,
               g.frame == FMM_FRAME_SHARED, ts);
    for (int k = 0; k < MAX_LEVELS; k++) c->sigT.s[k] = c->T.scale[k];
    int nts = c->T.num_slots;
    launch_gather_xyz(c->T.slots.as<int>(), nts, c->tgt_xyz.as<double>(), c->txyz.get<double>(3 * (size_t)std::max(nts, 1)),
                      ts);
    upload_leaves(c->T, c->h_tgt_leaves, c->tgt_leaves, ts);
}

// Both trees: the target side on the side stream by a host thread of its own (its own scratch)
// while this thread builds the source side; the main stream then waits for the side stream.
void stage_trees(fmm_ctx* c) {
    double tc[3], th[3];
    trees_begin(c, tc, th);
    std::exception_ptr tgt_error;
    std::thread tgt_thread([&] {
        try {
            build_target_side(c, tc, th, c->side);
            FMM_HIP(hipEventRecord(c->j_tgt, c->side));
        } catch (...) {
            tgt_error = std::current_exception();
        }
    });
    try {
        build_source_side(c);
    } catch (...) {
        tgt_thread.join();
        throw;
    }
    tgt_thread.join();
    if (tgt_error) std::rethrow_exception(tgt_error);
    FMM_HIP(hipStreamWaitEvent(c->stream, c->j_tgt, 0));
    c->rec(E_TREE);
    c->stage = ST_TREES;
}

void stage_leaves(fmm_ctx* c) {
    require(c->stage >= ST_TREES, FMM_ERR_STATE, "trees are not built");
    size_t n = (size_t)c->S.num_nodes * 3 * c->Lp;
    void* M = c->M.reserve(n * c->esize());
    c->leaves_split = false;
    // k_p2m writes its leaves whole, so only the other nodes are cleared (all of them for BEM)
    if (c->bem) FMM_HIP(hipMemsetAsync(M, 0, n * c->esize(), c->stream));
    else launch_zero_unwritten(c->S, c->Lp, c->esize(), M, c->stream);
    if (c->bem)   // Problem::BEM (multipole_solver.cpp:78-83 -> integration.cpp:295-353)
        launch_p2m_bem(c->f64(), c->src_leaves.as<int>(), (int)c->h_src_leaves.size(), c->S, c->sigS,
                       c->sxyz.as<double>(), c->sq.as<double>(), c->sbd.as<double>(), c->p, c->Lp, M, c->stream);
    else {
        // F16X3: the P2M also writes its leaves' columns of the M2L's split multipoles (the split
        // after the M2M then covers the other nodes only)
        F16Leaves split;
        c->leaves_split = c->prec == Precision::F16X3 && c->S.num_nodes > 0;
        if (c->leaves_split) {
            const int ncols = 3 * c->S.num_nodes, kp = m2l_f16_kpad(c->L);
            split.dm = c->dm.as<float>();
            split.Kp = kp;
            split.cscale = c->cscale.get<float>(ncols);
            split.hl = c->Mhl.get<_Float16>((size_t)2 * ncols * kp);
        }
        launch_p2m(c->f64(), c->src_leaves.as<int>(), (int)c->h_src_leaves.size(), c->S, c->sigS,
                   c->sxyz.as<double>(), c->sq.as<double>(), c->p, c->Lp, M, c->stream, split);
    }
    size_t nss = (size_t)std::max(c->S.num_slots, 1), rs = c->f64() ? sizeof(double) : sizeof(float);
    // origin = source root centre (node 0); the P2P kernel reads the same three doubles
    launch_pack_sources(c->f64(), c->S.num_slots, c->S.center.as<double>(), c->sxyz.as<double>(), c->sq.as<double>(),
                        c->nP0.reserve(4 * rs * nss), c->nP1.reserve(4 * rs * nss), c->nP2.reserve(rs * nss),
                        c->stream);
    c->rec(E_P2M);
    c->stage = ST_LEAVES;
}

// M2M of the non-matrix algorithms (MultipoleSolver::calclMultipoleExpansions, multipole_solver.cpp:
// 102-339): NoTranslation = P2M at every subdivided node from all its points (the slot ranges of
// a subtree are contiguous); ComplexTranslation, RealTranslation and Layers = the per-node
// translation formula, level by level, without operator tables
void upward_without_matrices(fmm_ctx* c, int alg) {
    const TreeDev& S = c->S;
    hipStream_t st = c->stream;
    std::vector<int> internal;
    std::vector<int> level_begin(1, 0);
    for (int k = 0; k + 1 < S.num_levels; k++) {
        for (int n = S.level_begin[k]; n < S.level_begin[k + 1]; n++)
            if (S.h_first_child[n] >= 0) internal.push_back(n);
        level_begin.push_back((int)internal.size());
    }
    if (internal.empty()) return;
    Staging& sg = c->up_stage;
    sg.begin();
    const size_t o = sg.put(internal);
    const int* d = (const int*)(sg.send(st) + o);
    if (alg == FMM_M2M_NO_TRANSLATION) {
        launch_p2m(c->f64(), d, (int)internal.size(), S, c->sigS, c->sxyz.as<double>(), c->sq.as<double>(), c->p, c->Lp,
                   c->M.ptr, st);
        return;
    }
    for (int k = S.num_levels - 2; k >= 0; k--)   // parents of level k, deepest first
        m2m_per_node(c->f64(), d + level_begin[k], level_begin[k + 1] - level_begin[k], S, c->sigS, c->p, c->Lp,
                     c->M.ptr, st);
}

// M2M: per level, 8 orientation classes of (child -> parent) pairs (multipole_solver.cpp:341-423)
void stage_upward(fmm_ctx* c, int alg = FMM_M2M_MATRICES) {
    require(c->stage >= ST_LEAVES, FMM_ERR_STATE, "leaf multipoles are not ready");
    require(alg >= FMM_M2M_NO_TRANSLATION && alg <= FMM_M2M_MATRICES, FMM_ERR_ARG, "unknown M2M algorithm");
    const TreeDev& S = c->S;
    int nlev = S.num_levels;
    OpTables& O = c->up_ops;
    O.nops = 8 * std::max(nlev - 1, 0);
    // operator per (level, orientation) from the first node's own offset (the reference's
    // choice); the nominal offsets (child - parent = +-half) for orientations a level lacks
    std::vector<double> nominal(3 * (size_t)std::max(O.nops, 1));
    for (int k = 1; k < nlev; k++)
        for (int o = 0; o < 8; o++)
            for (int a = 0; a < 3; a++) nominal[((k - 1) * 8 + o) * 3 + a] = SIGN[o][a] * S.half[k][a];
    // per level: many parents -> parent-major tiles, each parent column summing its 8 children
    // (one K segment per orientation operator; empty children hold zero multipoles), written
    // once without atomics; few parents -> the 8 orientations as parallel classes of
    // (child, parent) pairs with atomic accumulation (a long serial K loop would dominate)
    if (O.nops) {   // built for every algorithm: the adjoint applies them transposed
        Staging& sp = c->up_specs;
        sp.begin();
        const size_t o = sp.put(nominal);
        const double* d_nominal = (const double*)(sp.send(c->stream) + o);
        OpSpec* d_specs = O.specs.get<OpSpec>(O.nops);
        make_m2m_specs(S, c->sigS, d_nominal, d_specs, c->stream);
        build_tables(c, O, d_specs, c->stream);
    }
    if (alg != FMM_M2M_MATRICES) {
        upward_without_matrices(c, alg);
        c->rec(E_M2M);
        c->stage = ST_UP;
        return;
    }
    Staging& sg = c->up_stage;
    sg.begin();
    c->m2m_pass.assign(MAX_LEVELS, Pass());
    std::vector<char> parent_major(MAX_LEVELS, 0);
    for (int k = nlev - 1; k >= 1; k--) {
        std::vector<int> first, parent;
        for (int n = S.level_begin[k - 1]; n < S.level_begin[k]; n++) {
            int fc = S.h_first_child[n];
            if (fc >= 0) { first.push_back(fc); parent.push_back(n); }
        }
        Pass& P = c->m2m_pass[k];
        std::vector<GemmTile> tiles;
        if (3 * first.size() >= (size_t)M2M_PARENT_MAJOR_COLUMNS) {
            std::vector<int> cb = {0, (int)first.size()}, cop = {(k - 1) * 8};
            plan_tiles(cb, cop, tiles);
            stage_pass(sg, P, first, parent, tiles);
            parent_major[k] = 1;
        } else {
            std::vector<int> all_s, all_d, cb(1, 0), cop;
            for (int o = 0; o < 8; o++) {
                for (size_t i = 0; i < first.size(); i++)
                    if (nonempty_host(S, first[i] + o)) { all_s.push_back(first[i] + o); all_d.push_back(parent[i]); }
                cb.push_back((int)all_s.size());
                cop.push_back((k - 1) * 8 + o);
            }
            plan_tiles(cb, cop, tiles);
            stage_pass(sg, P, all_s, all_d, tiles);
        }
    }
    const char* dev = sg.send(c->stream);
    for (int k = nlev - 1; k >= 1; k--) {
        Pass& P = c->m2m_pass[k];
        resolve_pass(P, dev);
        if (!P.ntiles) continue;
        if (parent_major[k])
            class_gemm(c->pass_prec(), ops_ptr(c, O), O.lo.ptr, P.tiles, P.ntiles, P.src, P.dst, c->M.ptr, c->M.ptr, c->Lp,
                       EPI_ADD, c->stream, false, 8, TRI_LOWER);
        else
            run_pass(c, O, P, c->M.ptr, c->M.ptr, EPI_ATOMIC, TRI_LOWER);
    }
    c->rec(E_M2M);
    c->stage = ST_UP;
}

// Reference frame: every M2L and L2L translation vector is unique (bounding-box target
// root, float-rounded centres), so operators are built per pair, in chunks, and applied
// with the same GEMM kernels (one 3-column tile per pair).  A correctness path: the
// shared frame is the fast one.
void run_per_pair(fmm_ctx* c, int kind, const int* d_src, const int* d_dst, int npairs, const TreeDev& Tsrc,
                  const TreeDev& Tdst, const ScaleTable& sg_src, const ScaleTable& sg_dst, const void* src_exp,
                  void* dst_exp, bool atomic) {
    hipStream_t st = c->stream;
    OpTables& O = c->down_ops;
    const size_t per = (size_t)c->Lp * c->Lp * sizeof(double);
    int chunk = (int)std::max<size_t>(1, ((size_t)256 << 20) / per);
    std::vector<GemmTile> tiles;
    for (int b = 0; b < npairs; b += chunk) {
        int n = std::min(chunk, npairs - b);
        OpSpec* sp = O.specs.get<OpSpec>(n);
        make_pair_specs(d_src + b, d_dst + b, n, Tsrc.center.as<double>(), Tdst.center.as<double>(),
                        Tsrc.level.as<int>(), Tdst.level.as<int>(), sg_src, sg_dst, kind, c->cfg.m2l_full, sp, st);
        O.nops = n;
        build_tables(c, O, sp, st);
        tiles.resize(n);
        for (int i = 0; i < n; i++) tiles[i] = GemmTile{i, 3 * (b + i), 3};
        GemmTile* dt = c->pp_tiles.get<GemmTile>(n);
        FMM_HIP(hipMemcpyAsync(dt, tiles.data(), n * sizeof(GemmTile), hipMemcpyHostToDevice, st));
        class_gemm(c->pass_prec(), ops_ptr(c, O), O.lo.ptr, dt, n, d_src, d_dst, src_exp, dst_exp, c->Lp,
                   atomic ? EPI_ATOMIC : EPI_ADD, st, false, 1, kind == OPK_L2L ? TRI_UPPER : TRI_NONE);
        FMM_HIP(hipStreamSynchronize(st));   // the chunk buffers are reused
    }
}

void downward_per_pair(fmm_ctx* c) {
    hipStream_t st = c->stream;
    const TreeDev& S = c->S;
    const TreeDev& T = c->T;
    size_t n = (size_t)T.num_nodes * 3 * c->Lp;
    void* Lx = c->Lx.reserve(n * c->esize());
    FMM_HIP(hipMemsetAsync(Lx, 0, n * c->esize(), st));
    c->rec(E_TABLES);
    run_per_pair(c, OPK_M2L, c->lists.far_s.as<int>(), c->lists.far_t.as<int>(), c->lists.num_far, S, T, c->sigS,
                 c->sigT, c->M.ptr, Lx, true);
    c->rec(E_M2L);
    // L2L with each child's exact translation (calculation_point_octree.cpp:277-299)
    Staging& sg = c->down_stage;
    sg.begin();
    c->l2l_pass.assign(MAX_LEVELS, Pass());
    for (int k = 1; k < T.num_levels; k++) {
        std::vector<int> src, dst;
        for (int nd = T.level_begin[k - 1]; nd < T.level_begin[k]; nd++) {
            int fc = T.h_first_child[nd];
            if (fc < 0) continue;
            for (int o = 0; o < 8; o++)
                if (nonempty_host(T, fc + o)) { src.push_back(nd); dst.push_back(fc + o); }
        }
        stage_pass(sg, c->l2l_pass[k], src, dst, std::vector<GemmTile>());
    }
    const char* dev = sg.send(st);
    for (int k = 1; k < T.num_levels; k++) {
        Pass& P = c->l2l_pass[k];
        resolve_pass(P, dev);
        P.ntiles = 0;   // per-pair operators: this pass has no shared tiles (no adjoint)
        if (P.npairs)
            run_per_pair(c, OPK_L2L, P.src, P.dst, P.npairs, T, T, c->sigT, c->sigT, Lx, Lx, false);
    }
    c->down_ops.nops = 0;
    c->m2l_pass.ntiles = 0;
    c->rec(E_L2L);
    c->stage = ST_DOWN;
}

// interaction lists (and, in the shared frame, far pairs grouped by target node for the
// atomic-free M2L accumulation) on stream ls; they depend on the trees only
// far_ready (may be null): recorded on ls once the far pairs and classes are complete; then
// after_far (may be empty) is called, and the close pairs and the target grouping follow, then
// E_LISTS.  The caller checks that the trees are built (this may run on a thread of its own).
void stage_lists(fmm_ctx* c, hipStream_t ls, hipEvent_t far_ready, const std::function<void()>& after_far = {}) {
    const bool shared = c->cfg.frame == FMM_FRAME_SHARED;
    c->rec_on(E_LISTS0, ls);
    // the pair index (a sort of the far pairs by target) is queued as soon as the far pairs are,
    // ahead of the close lists
    build_lists(c->lists, c->lscr, c->S, c->T, c->cfg.mac_scale, shared, ls, far_ready, [&] {
        if (shared) build_pair_index(c->lists.far_t.as<int>(), c->lists.num_far, c->T.num_nodes, c->m2l_ix, ls);
        if (after_far) after_far();
    });
    c->rec_on(E_LISTS, ls);
    c->lists_ready = true;
}

// Shared frame, downward pass part 1 on stream st: operator tables (L2L per level and
// orientation, one M2L operator per class), the M2L tiles and the cleared local expansions.
// Needs the far pairs and classes only (not the multipoles); the fused evaluation queues it on
// the main stream after the M2M and the column split, once the far classes are known.
void translate_tables(fmm_ctx* c, hipStream_t st) {
    const TreeDev& S = c->S;
    const TreeDev& T = c->T;
    c->rec_on(E_DOWN0, st);
    // operator tables: L2L per (level, orientation), then one M2L operator per class
    OpTables& O = c->down_ops;
    int nlev = T.num_levels;
    int n_l2l = 8 * std::max(nlev - 1, 0);
    int nc = c->lists.num_classes;
    O.nops = n_l2l + nc;
    std::vector<OpSpec> specs(O.nops);   // the M2L specs are written on the device
    for (int k = 1; k < nlev; k++)
        for (int o = 0; o < 8; o++) {
            OpSpec& s = specs[(k - 1) * 8 + o];
            for (int a = 0; a < 3; a++) s.t[a] = -SIGN[o][a] * T.half[k][a];   // parent - child
            s.sig_in = T.scale[k - 1];
            s.sig_out = T.scale[k];
            s.kind = OPK_L2L;
            s.m2l_full = 0;
        }
    Staging& sp = c->m2l_specs;
    sp.begin();
    const size_t o_specs = sp.put(specs);
    OpSpec* d = (OpSpec*)(sp.send(st) + o_specs);
    make_m2l_specs(c->lists.class_begin.as<int>(), nc, c->lists.far_t.as<int>(), c->lists.far_s.as<int>(),
                   S.center.as<double>(), T.center.as<double>(), S.level.as<int>(), T.level.as<int>(), c->sigS, c->sigT,
                   c->cfg.m2l_full, d + n_l2l, st);
    c->down_l2l = n_l2l;
    c->down_specs = d;
    if (c->prec == Precision::F16X3) {
        build_tables(c, O, d, st, n_l2l);   // L2L in FP32
        if (nc) {
            F16Split f;
            f.R = m2l_f16_rows(c->L);
            f.Kp = m2l_f16_kpad(c->L);
            f.hi = c->m2l_hi.get<_Float16>((size_t)nc * f.R * f.Kp);
            f.lo = c->m2l_lo.get<_Float16>((size_t)nc * f.R * f.Kp);
            f.rinv = c->m2l_rinv.get<float>((size_t)nc * f.R);
            f.dm = c->dm.as<float>();
            make_m2l_f16(d + n_l2l, nc, c->p, f, c->m2l_work.reserve(m2l_f16_work_bytes(nc, c->p)), st);
        }
    } else if (O.nops) {
        build_tables(c, O, d, st);
    }
    // the M2L tiles (planned while the tables are built)
    Pass& P = c->m2l_pass;
    {
        std::vector<int> cop(nc);
        for (int i = 0; i < nc; i++) cop[i] = n_l2l + i;
        std::vector<GemmTile> tiles;
        plan_tiles(c->lists.h_class_begin, cop, tiles, m2l_tile_cols(c));
        Staging& sg = c->m2l_stage;
        sg.begin();
        const size_t o_tiles = sg.put(tiles);
        P.npairs = c->lists.num_far;
        P.ntiles = (int)tiles.size();
        P.width = m2l_tile_cols(c);
        P.tiles = (const GemmTile*)(sg.send(st) + o_tiles);
    }
    // the pair reduction writes every local expansion whole; cleared here only when it will not run
    size_t n = (size_t)T.num_nodes * 3 * c->Lp;
    void* Lx = c->Lx.reserve(n * c->esize());
    if (!P.ntiles) FMM_HIP(hipMemsetAsync(Lx, 0, n * c->esize(), st));
    c->rec_on(E_TABLES, st);
}

// F16X3, stream A after the multipoles: the B side of the FP16 split of the M2L (column scales
// and hi / lo parts of the source nodes' multipoles, except the leaves' the P2M already split);
// needs neither the lists nor the tables.
void split_for_gemm(fmm_ctx* c) {
    const int ncols = 3 * c->S.num_nodes, kp = m2l_f16_kpad(c->L);
    const bool leaves = c->leaves_split;   // the P2M wrote its leaves' columns
    split_columns(c->M.as<float>(), ncols, c->L, c->Lp, kp, c->dm.as<float>(), c->cscale.get<float>(ncols),
                  c->Mhl.get<_Float16>((size_t)2 * ncols * kp), c->stream,
                  leaves ? c->S.first_child.as<int>() : nullptr, leaves ? c->S.npts.as<int>() : nullptr);
}

// Part 2 (stream A, after the tables and the multipoles): the M2L of all classes in one launch
// (fast_multipole_solver.cpp:258-276) writing one slab per (pair, component); each target node
// then sums its pairs' slabs in pair order (part 3: no atomics, deterministic).  E_GEMM0 ..
// E_M2L_GEMM brackets the GEMM alone.  split_done: split_for_gemm has run on stream A.
void translate_gemm(fmm_ctx* c, hipEvent_t gemm_after = nullptr, bool split_done = false) {
    hipStream_t st = c->stream;
    const OpTables& O = c->down_ops;
    const Pass& P = c->m2l_pass;
    const bool f16 = c->prec == Precision::F16X3;
    if (f16 && P.ntiles && !split_done) split_for_gemm(c);
    if (gemm_after) FMM_HIP(hipStreamWaitEvent(st, gemm_after, 0));
    c->rec(E_GEMM0);
    if (P.ntiles) {
        const int R = m2l_slab_rows(c->prec, c->L, c->Lp);
        void* slab = c->m2l_slab.reserve((size_t)3 * P.npairs * R * c->esize());
        if (f16)
            m2l_gemm_f16x3(c->m2l_hi.as<_Float16>(), c->m2l_lo.as<_Float16>(), c->m2l_rinv.as<float>(), c->down_l2l,
                           P.tiles, P.ntiles, m2l_tile_cols(c), c->lists.far_s.as<int>(), c->Mhl.as<_Float16>(),
                           c->cscale.as<float>(), (float*)slab, c->L, c->queues.get<int>(16), st);
        else
            m2l_gemm(c->prec, ops_ptr(c, O), O.lo.ptr, P.tiles, P.ntiles, c->lists.far_s.as<int>(), c->M.ptr, slab,
                     c->L, c->Lp, st);
    }
    c->rec(E_M2L_GEMM);
    FMM_HIP(hipEventRecord(c->j_gemm, st));
}

// Part 3: per-target reduction of the slab (needs the pair index), then L2L
void translate_finish(fmm_ctx* c) {
    hipStream_t st = c->stream;
    const TreeDev& T = c->T;
    OpTables& O = c->down_ops;
    const int nlev = T.num_levels;
    void* Lx = c->Lx.ptr;
    if (c->m2l_pass.ntiles) {
        reduce_pairs(c->f64(), c->m2l_ix, c->m2l_slab.ptr, m2l_slab_rows(c->prec, c->L, c->Lp), c->Lp, Lx, st);
    }
    c->rec(E_M2L);
    // L2L top-down (calculation_point_octree.cpp:277-299)
    Staging& sg = c->down_stage;
    sg.begin();
    c->l2l_pass.assign(MAX_LEVELS, Pass());
    for (int k = 1; k < nlev; k++) {
        std::vector<int> src[8], dst[8];
        for (int nd = T.level_begin[k - 1]; nd < T.level_begin[k]; nd++) {
            int fc = T.h_first_child[nd];
            if (fc < 0) continue;
            for (int o = 0; o < 8; o++)
                if (nonempty_host(T, fc + o)) { src[o].push_back(nd); dst[o].push_back(fc + o); }
        }
        std::vector<int> all_s, all_d, cb(1, 0), cop;
        for (int o = 0; o < 8; o++) {
            all_s.insert(all_s.end(), src[o].begin(), src[o].end());
            all_d.insert(all_d.end(), dst[o].begin(), dst[o].end());
            cb.push_back((int)all_s.size());
            cop.push_back((k - 1) * 8 + o);
        }
        std::vector<GemmTile> tiles;
        plan_tiles(cb, cop, tiles);
        stage_pass(sg, c->l2l_pass[k], all_s, all_d, tiles);
    }
    const char* dev = sg.send(st);
    for (int k = 1; k < nlev; k++) {
        resolve_pass(c->l2l_pass[k], dev);
        run_pass(c, O, c->l2l_pass[k], Lx, Lx, EPI_ADD, TRI_UPPER);
    }
    c->rec(E_L2L);
    c->stage = ST_DOWN;
}

// M2L and L2L (the lists are ready on c->stream's timeline)
void stage_translate(fmm_ctx* c) {
    require(c->stage >= ST_UP, FMM_ERR_STATE, "Multipoles are not ready!");
    require(c->lists_ready, FMM_ERR_STATE, "interaction lists are not built");
    if (c->cfg.frame != FMM_FRAME_SHARED) {
        c->rec(E_DOWN0);
        downward_per_pair(c);
        return;
    }
    translate_tables(c, c->stream);
    translate_gemm(c);
    translate_finish(c);
}

void stage_downward(fmm_ctx* c) {
    require(c->stage >= ST_UP, FMM_ERR_STATE, "Multipoles are not ready!");
    stage_lists(c, c->stream, c->ev_ok ? c->ev[E_FAR] : nullptr);   // ST_UP implies the trees
    stage_translate(c);
}

__global__ void k_fill_nan(double* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = __longlong_as_double(0x7ff8000000000000ll);
}

// near field on stream ps (needs the lists and the packed sources)
void stage_p2p(fmm_ctx* c, hipStream_t ps, bool queue = false) {
    int nts = c->T.num_slots;
    double* Bnear = c->Bnear.get<double>(3 * (size_t)std::max(nts, 1));
    c->rec_on(E_P2P0, ps);
    if (c->cfg.near_field == FMM_NEAR_TREECODE) {
        // the reference's close list: M2P walks from each close node, failing leaves dropped
        FMM_HIP(hipMemsetAsync(Bnear, 0, 3 * (size_t)std::max(nts, 1) * sizeof(double), ps));
        launch_m2p_close(c->f64(), c->S, c->sigS, c->M.ptr, c->p, c->Lp, c->tgt_leaves.as<int>(),
                         (int)c->h_tgt_leaves.size(), c->T, c->lists.close_begin.as<int>(), c->lists.close_end.as<int>(),
                         c->lists.close_s.as<int>(), c->txyz.as<double>(), Bnear, nullptr, ps);
    } else if (c->cfg.near_field && queue) {
        // one persistent workgroup per CU beside the M2L GEMM (see run_fused)
        launch_p2p_queue(c->f64(), c->tgt_leaves.as<int>(), (int)c->h_tgt_leaves.size(), c->T, c->S,
                         c->lists.close_begin.as<int>(), c->lists.close_end.as<int>(), c->lists.close_s.as<int>(),
                         c->nP0.ptr, c->nP1.ptr, c->nP2.ptr, c->txyz.as<double>(), Bnear, c->queues.get<int>(16) + 8,
                         num_cus(), ps);
    } else if (c->cfg.near_field) {
        launch_p2p(c->f64(), c->tgt_leaves.as<int>(), (int)c->h_tgt_leaves.size(), c->T, c->S, c->lists.close_begin.as<int>(),
                   c->lists.close_end.as<int>(), c->lists.close_s.as<int>(), c->nP0.ptr, c->nP1.ptr, c->nP2.ptr,
                   c->txyz.as<double>(), Bnear, ps);
    } else {
        FMM_HIP(hipMemsetAsync(Bnear, 0, 3 * (size_t)std::max(nts, 1) * sizeof(double), ps));
    }
    c->rec_on(E_P2P, ps);
}

void stage_l2p(fmm_ctx* c) {
    require(c->stage >= ST_DOWN, FMM_ERR_STATE, "local expansions are not ready");
    int nts = c->T.num_slots;
    double* Bfar = c->Bfar.get<double>(3 * (size_t)std::max(nts, 1));
    launch_l2p(c->f64(), c->tgt_leaves.as<int>(), (int)c->h_tgt_leaves.size(), c->T, c->sigT, c->txyz.as<double>(),
               c->Lx.ptr, c->p, c->Lp, Bfar, c->stream);
    c->rec(E_L2P);
}

void stage_assemble(fmm_ctx* c, double current, double* Bdev) {
    hipStream_t st = c->stream;
    int nts = c->T.num_slots;
    size_t nout = 3 * (size_t)c->nt;
    if (nout) k_fill_nan<<<(unsigned)((nout + 255) / 256), 256, 0, st>>>(Bdev, nout);
    c->current = current;
    launch_assemble(c->T.slots.as<int>(), nts, c->Bfar.as<double>(), c->Bnear.as<double>(),
                    current * REF_MU0 / (4.0 * REF_PI), Bdev, st);
    c->rec(E_END);
    c->stage = ST_EVAL;
}

void stage_evaluate(fmm_ctx* c, double current, double* Bdev) {
    require(c->stage >= ST_DOWN, FMM_ERR_STATE, "local expansions are not ready");
    stage_l2p(c);
    stage_p2p(c, c->stream);
    stage_assemble(c, current, Bdev);
}

// calcA(current) (fast_multipole_solver.cpp:195-212): A = current / (4 pi) (sum L R(x - c) over
// the target's leaf + the close list), input order; targets outside the target tree get NaN.  For
// BEM sources this is calcBEM (:233-250) with current = 1.
void stage_evaluate_A(fmm_ctx* c, double current, double* Adev) {
    require(c->stage >= ST_DOWN, FMM_ERR_STATE, "local expansions are not ready");
    hipStream_t st = c->stream;
    const int nts = c->T.num_slots, nl = (int)c->h_tgt_leaves.size();
    const size_t n3 = 3 * (size_t)std::max(nts, 1);
    double* Af = c->Afar.get<double>(n3);
    double* An = c->Anear.get<double>(n3);
    launch_l2p_A(c->f64(), c->tgt_leaves.as<int>(), nl, c->T, c->sigT, c->txyz.as<double>(), c->Lx.ptr, c->p, c->Lp,
                 Af, st);
    FMM_HIP(hipMemsetAsync(An, 0, n3 * sizeof(double), st));
    if (c->cfg.near_field == FMM_NEAR_TREECODE)
        launch_m2p_close(c->f64(), c->S, c->sigS, c->M.ptr, c->p, c->Lp, c->tgt_leaves.as<int>(), nl, c->T,
                         c->lists.close_begin.as<int>(), c->lists.close_end.as<int>(), c->lists.close_s.as<int>(),
                         c->txyz.as<double>(), nullptr, An, st);
    else if (c->cfg.near_field && c->bem)
        launch_p2p_bem(c->tgt_leaves.as<int>(), nl, c->T, c->S, c->lists.close_begin.as<int>(),
                       c->lists.close_end.as<int>(), c->lists.close_s.as<int>(), c->sxyz.as<double>(),
                       c->sq.as<double>(), c->sbd.as<double>(), c->txyz.as<double>(), An, st);
    else if (c->cfg.near_field)
        launch_p2p_A(c->f64(), c->tgt_leaves.as<int>(), nl, c->T, c->S, c->lists.close_begin.as<int>(),
                     c->lists.close_end.as<int>(), c->lists.close_s.as<int>(), c->nP0.ptr, c->nP1.ptr, c->nP2.ptr,
                     c->txyz.as<double>(), An, st);
    const size_t nout = 3 * (size_t)c->nt;
    if (nout) k_fill_nan<<<(unsigned)((nout + 255) / 256), 256, 0, st>>>(Adev, nout);
    launch_assemble(c->T.slots.as<int>(), nts, Af, An, current / (4.0 * REF_PI), Adev, st);
    FMM_HIP(hipGetLastError());
}

// the single-point treecode (multipole_solver.cpp:686-700) at m device points: B scaled by
// current MU0 / (4 pi), A by current / (4 pi); either may be null
void stage_treecode(fmm_ctx* c, double current, size_t m, const double* xyz, double* B, double* A) {
    require(c->stage >= ST_UP, FMM_ERR_STATE, "Multipoles are not ready!");
    require(m < (size_t)INT32_MAX / 3, FMM_ERR_LIMIT, "too many points");
    require(m == 0 || xyz, FMM_ERR_ARG, "null point array");
    launch_treecode_points(c->f64(), c->S, c->sigS, c->M.ptr, c->p, c->Lp, xyz, (int)m,
                           current * REF_MU0 / (4.0 * REF_PI), current / (4.0 * REF_PI), B, A, c->stream);
}

// FP64 direct sum of B (pot: A) at m device targets into out, partial sums in part (nsplit *
// m * 3 doubles)
void direct_sum(fmm_ctx* c, bool pot, double current, size_t m, const double* xyz, double* out, double* part,
                int nsplit) {
    const double scale = pot ? current / (4.0 * REF_PI) : current * REF_MU0 / (4.0 * REF_PI);
    launch_direct(pot, c->sxyz.as<double>(), c->sq.as<double>(), c->S.num_slots, xyz, (int)m, scale, out, part,
                  nsplit, c->stream);
}

// The whole evaluation with the side stream: the lists are built (on a host thread, since
// their compaction reads counts back) while P2M and M2M run, and the near field runs beside the
// M2L GEMM (persistent F16X3 GEMM) or after it, beside the memory-bound slab reduction, L2L and
// L2P; the streams join before the assembly.  The reference frame's per-pair M2L has no GEMM
// event split and runs staged.
void run_fused(fmm_ctx* c, double current, double* Bdev) {
    if (c->cfg.frame != FMM_FRAME_SHARED) {
        stage_trees(c);
        stage_leaves(c);
        stage_upward(c);
        stage_downward(c);
        stage_evaluate(c, current, Bdev);
        return;
    }
    hipStream_t A = c->stream, Bs = c->side;
    // Side thread (stream Bs): target tree, targets in slot order, then -- once the source tree
    // is built -- the far lists (E_FAR), the close lists and the pair index (j_lists).  This
    // thread (stream A): source tree, charges, P2M, M2M, then -- once the far classes are known
    // on the host -- the operator tables and the M2L GEMM.  Neither thread blocks on the other's
    // GPU work except through events; the side thread only waits for the counts it reads back.
    double tc[3], th[3];
    trees_begin(c, tc, th);
    std::mutex mu;
    std::condition_variable cv;
    int src_state = 0, far_state = 0, lists_state = 0, gemm_state = 0;   // 0 pending, 1 ready, 2 failed
    // F16X3 with the direct near field and the persistent M2L GEMM (p = 14..18): the P2P runs as a
    // persistent grid of one workgroup per CU beside the GEMM (whose register budget leaves it
    // room), launched when the GEMM's inputs are complete (E_GEMM0, after the column split).
    // Launched earlier, beside the M2M and the tables, it left the GEMM up to 1.4 ms slower --
    // most likely because workgroups placed on a busy chip can leave a CU without room for a
    // GEMM workgroup.  Otherwise the P2P follows the GEMM (beside the GEMMs of one workgroup per
    // tile it only slows them).
    const bool p2p_beside = c->prec == Precision::F16X3 && c->cfg.near_field == FMM_NEAR_P2P &&
                            m2l_tile_cols(c) == 128 && m2l_f16x3_persistent(c->L);
    std::exception_ptr side_error;
    auto post = [&](int& state, int v) {
        {
            std::lock_guard<std::mutex> lk(mu);
            state = v;
        }
        cv.notify_all();
    };
    std::thread side_thread([&] {
        try {
            build_target_side(c, tc, th, Bs);
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return src_state != 0; });
            }
            if (src_state == 1) {
                FMM_HIP(hipStreamWaitEvent(Bs, c->j_src, 0));   // the source tree on stream A
                stage_lists(c, Bs, c->ev[E_FAR], [&] { post(far_state, 1); });   // E_FAR also joins
                FMM_HIP(hipEventRecord(c->j_lists, Bs));
                post(lists_state, 1);
                if (p2p_beside) {
                    {
                        std::unique_lock<std::mutex> lk(mu);
                        cv.wait(lk, [&] { return gemm_state != 0; });
                    }
                    if (gemm_state == 1) {   // E_GEMM0 of this evaluation is recorded (it follows j_pack)
                        FMM_HIP(hipStreamWaitEvent(Bs, c->ev[E_GEMM0], 0));
                        stage_p2p(c, Bs, true);
                        FMM_HIP(hipEventRecord(c->j_p2p, Bs));
                    }
                }
            }
        } catch (...) {
            side_error = std::current_exception();
        }
        if (far_state == 0) post(far_state, 2);
        if (lists_state == 0) post(lists_state, 2);
    });
    try {
        build_source_side(c);
        FMM_HIP(hipEventRecord(c->j_src, A));
        post(src_state, 1);
        c->rec(E_TREE);
        c->stage = ST_TREES;
        stage_leaves(c);
        FMM_HIP(hipEventRecord(c->j_pack, A));   // packed near-field sources, for the P2P
        stage_upward(c);
        // the column split needs only the multipoles: queued before any wait for the side thread,
        // so that it follows the M2M at once (the tables, which need the far classes, follow it)
        const bool split_early = c->prec == Precision::F16X3 && c->S.num_nodes > 0;
        if (split_early) split_for_gemm(c);
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return far_state != 0; });
        }
        if (far_state == 1) {
            FMM_HIP(hipStreamWaitEvent(A, c->ev[E_FAR], 0));   // far pairs and classes
            translate_tables(c, A);
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return lists_state != 0; });
            }
            // the GEMM starts after the lists' last kernels (the pair index sort): beside the
            // persistent GEMM a sort whose blocks wait on each other crawled (17 ms instead of
            // 0.1), holding back the near field queued after it; normally the lists end first
            translate_gemm(c, lists_state == 1 ? c->j_lists : nullptr, split_early);
            post(gemm_state, 1);
        } else {
            post(gemm_state, 2);
        }
    } catch (...) {
        if (src_state == 0) post(src_state, 2);
        if (gemm_state == 0) post(gemm_state, 2);
        side_thread.join();
        throw;
    }
    side_thread.join();
    if (side_error) std::rethrow_exception(side_error);
    FMM_HIP(hipStreamWaitEvent(A, c->j_lists, 0));    // close lists, pair index
    translate_finish(c);
    if (!p2p_beside) {
        FMM_HIP(hipStreamWaitEvent(Bs, c->j_pack, 0));
        FMM_HIP(hipStreamWaitEvent(Bs, c->j_gemm, 0));
        stage_p2p(c, Bs);
        FMM_HIP(hipEventRecord(c->j_p2p, Bs));
    }
    stage_l2p(c);
    FMM_HIP(hipStreamWaitEvent(A, c->j_p2p, 0));
    stage_assemble(c, current, Bdev);
}

void set_sources_common(fmm_ctx* c, size_t n, const double* xyz, const double* w, const double* J, hipMemcpyKind kind) {
    require(n < (size_t)INT32_MAX / 3, FMM_ERR_LIMIT, "too many sources");
    if (kind == hipMemcpyDeviceToDevice) join_default_stream(c);
    require(n == 0 || (xyz && w), FMM_ERR_ARG, "null source arrays");
    c->ns = (int)n;
    c->have_J = J != nullptr;
    c->bem = false;
    size_t nn = std::max<size_t>(n, 1);
    if (n) {
        FMM_HIP(hipMemcpyAsync(c->src_xyz.get<double>(3 * nn), xyz, 3 * n * sizeof(double), kind, c->stream));
        FMM_HIP(hipMemcpyAsync(c->src_w.get<double>(nn), w, n * sizeof(double), kind, c->stream));
        if (J) FMM_HIP(hipMemcpyAsync(c->src_J.get<double>(3 * nn), J, 3 * n * sizeof(double), kind, c->stream));
    } else {
        c->src_xyz.get<double>(3);
        c->src_w.get<double>(1);
    }
    FMM_HIP(hipStreamSynchronize(c->stream));
    c->stage = ST_EMPTY;
}

// Problem::BEM sources: position, weight, field B and normal n per point (BEMQuadrature)
void set_bem_sources_common(fmm_ctx* c, size_t n, const double* xyz, const double* w, const double* Bv,
                            const double* nrm, hipMemcpyKind kind) {
    require(n == 0 || (Bv && nrm), FMM_ERR_ARG, "null BEM source arrays");
    set_sources_common(c, n, xyz, w, nullptr, kind);
    c->bem = true;
    size_t nn = std::max<size_t>(n, 1);
    double* dB = c->src_B.get<double>(3 * nn);
    double* dN = c->src_N.get<double>(3 * nn);
    if (n) {
        FMM_HIP(hipMemcpyAsync(dB, Bv, 3 * n * sizeof(double), kind, c->stream));
        FMM_HIP(hipMemcpyAsync(dN, nrm, 3 * n * sizeof(double), kind, c->stream));
    }
    FMM_HIP(hipStreamSynchronize(c->stream));
}

void set_targets_common(fmm_ctx* c, size_t m, const double* xyz, hipMemcpyKind kind) {
    require(m < (size_t)INT32_MAX / 3, FMM_ERR_LIMIT, "too many targets");
    if (kind == hipMemcpyDeviceToDevice) join_default_stream(c);
    require(m == 0 || xyz, FMM_ERR_ARG, "null target array");
    c->nt = (int)m;
    size_t mm = std::max<size_t>(m, 1);
    if (m) FMM_HIP(hipMemcpyAsync(c->tgt_xyz.get<double>(3 * mm), xyz, 3 * m * sizeof(double), kind, c->stream));
    else c->tgt_xyz.get<double>(3);
    // target extent, once per target set (math::max / math::min, math.cpp:60-84)
    std::vector<double> h;
    const double* P = xyz;
    if (kind == hipMemcpyDeviceToDevice && m) {
        h.resize(3 * m);
        FMM_HIP(hipMemcpyAsync(h.data(), xyz, 3 * m * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        P = h.data();
    }
    FMM_HIP(hipStreamSynchronize(c->stream));
    for (int a = 0; a < 3; a++) {
        double mx = -1e+16, mn = 1e+16;
        for (size_t i = 0; i < m; i++) {
            double v = P[3 * i + a];
            if (v > mx) mx = v;
            if (v < mn) mn = v;
        }
        c->tmin[a] = mn;
        c->tmax[a] = mx;
    }
    c->stage = ST_EMPTY;
}

}  // namespace

extern "C" {

void fmm_config_init(fmm_config* g) {
    if (!g) return;
    std::memset(g, 0, sizeof(*g));
    g->order = 10;
    g->source_leaf_capacity = 1000;   // reference defaults (fast_multipole_solver.hpp:29-30)
    g->target_leaf_capacity = 100;
    g->precision = FMM_PREC_FP64;
    g->frame = FMM_FRAME_REFERENCE;   // the reference's own trees and output order
    g->m2l_full = 1;
    g->max_depth = 21;
    g->near_field = 1;
    g->mac_scale = 1.0;
    for (int a = 0; a < 3; a++) { g->source_root_center[a] = 0.0; g->source_root_half[a] = 3.0; }
}

int fmm_create(fmm_ctx** out, const fmm_config* cfg) {
    if (!out || !cfg) return FMM_ERR_ARG;
    *out = nullptr;
    if (cfg->order < 1 || cfg->order > MAX_ORDER) return FMM_ERR_ARG;
    if (cfg->source_leaf_capacity < 2 || cfg->target_leaf_capacity < 1) return FMM_ERR_ARG;
    if (cfg->precision < 0 || cfg->precision > 3 || cfg->frame < 0 || cfg->frame > 1) return FMM_ERR_ARG;
    if (!(cfg->mac_scale >= 1.0)) return FMM_ERR_ARG;
    if (cfg->near_field < 0 || cfg->near_field > FMM_NEAR_TREECODE) return FMM_ERR_ARG;
    fmm_ctx* c = new fmm_ctx();
    c->cfg = *cfg;
    if (c->cfg.max_depth <= 0 || c->cfg.max_depth > MAX_LEVELS - 1) c->cfg.max_depth = MAX_LEVELS - 1;
    c->p = cfg->order;
    c->L = nlen(c->p);
    c->Lp = pad_to(c->L, 64);
    c->prec = (Precision)cfg->precision;
    int rc = guard(c, [&] {
        int ndev = 0;
        FMM_HIP(hipGetDeviceCount(&ndev));
        if (ndev < 1) throw Error(FMM_ERR_HIP, "no HIP device");
        FMM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
        c->scratch = scratch_on_current_device();
        c->queues.get<int>(16);   // allocated once: both host threads of an evaluation use it
        for (int i = 0; i < E_COUNT; i++) FMM_HIP(hipEventCreate(&c->ev[i]));
        c->ev_ok = true;
        FMM_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));   // default priority
        if (c->prec == Precision::F16X3) {
            const int kp = m2l_f16_kpad(c->L);
            std::vector<float> h = coefficient_scales(c->p, kp);
            FMM_HIP(hipMemcpy(c->dm.get<float>(kp), h.data(), kp * sizeof(float), hipMemcpyHostToDevice));
        }
        FMM_HIP(hipEventCreateWithFlags(&c->j_lists, hipEventDisableTiming));
        FMM_HIP(hipEventCreateWithFlags(&c->j_pack, hipEventDisableTiming));
        FMM_HIP(hipEventCreateWithFlags(&c->j_p2p, hipEventDisableTiming));
        FMM_HIP(hipEventCreateWithFlags(&c->j_gemm, hipEventDisableTiming));
        FMM_HIP(hipEventCreateWithFlags(&c->j_start, hipEventDisableTiming));
        FMM_HIP(hipEventCreateWithFlags(&c->j_tables, hipEventDisableTiming));
        FMM_HIP(hipEventCreateWithFlags(&c->j_src, hipEventDisableTiming));
        FMM_HIP(hipEventCreateWithFlags(&c->j_tgt, hipEventDisableTiming));
        FMM_HIP(hipEventCreateWithFlags(&c->j_caller, hipEventDisableTiming));
    });
    if (rc != FMM_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return FMM_OK;
}

void fmm_destroy(fmm_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ev_ok)
        for (int i = 0; i < E_COUNT; i++) (void)hipEventDestroy(c->ev[i]);
    for (hipEvent_t e : {c->j_lists, c->j_pack, c->j_p2p, c->j_gemm, c->j_start, c->j_tables, c->j_src, c->j_tgt,
                         c->j_caller})
        if (e) (void)hipEventDestroy(e);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* fmm_last_error(const fmm_ctx* c) { return c ? c->err.c_str() : "null context"; }

int fmm_set_stream(fmm_ctx* c, void* s) {
    return guard(c, [&] {
        FMM_HIP(hipStreamSynchronize(c->stream));
        if (s) {
            if (c->own_stream) FMM_HIP(hipStreamDestroy(c->stream));
            c->own_stream = false;
            c->stream = (hipStream_t)s;
        } else if (!c->own_stream) {
            // back to a stream of the context's own (the caller's stream is no longer used)
            hipStream_t own = nullptr;
            FMM_HIP(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
            c->stream = own;
            c->own_stream = true;
        }
    });
}

int fmm_synchronize(fmm_ctx* c) { return guard(c, [&] { FMM_HIP(hipStreamSynchronize(c->stream)); }); }

int fmm_set_sources(fmm_ctx* c, size_t n, const double* xyz, const double* w, const double* J) {
    return guard(c, [&] { set_sources_common(c, n, xyz, w, J, hipMemcpyHostToDevice); });
}
int fmm_set_targets(fmm_ctx* c, size_t m, const double* xyz) {
    return guard(c, [&] { set_targets_common(c, m, xyz, hipMemcpyHostToDevice); });
}
int fmm_set_sources_dev(fmm_ctx* c, size_t n, const double* xyz, const double* w, const double* J) {
    return guard(c, [&] { set_sources_common(c, n, xyz, w, J, hipMemcpyDeviceToDevice); });
}
int fmm_set_targets_dev(fmm_ctx* c, size_t m, const double* xyz) {
    return guard(c, [&] { set_targets_common(c, m, xyz, hipMemcpyDeviceToDevice); });
}
int fmm_set_bem_sources(fmm_ctx* c, size_t n, const double* xyz, const double* w, const double* B,
                        const double* normal) {
    return guard(c, [&] { set_bem_sources_common(c, n, xyz, w, B, normal, hipMemcpyHostToDevice); });
}
int fmm_set_bem_sources_dev(fmm_ctx* c, size_t n, const double* xyz, const double* w, const double* B,
                            const double* normal) {
    return guard(c, [&] { set_bem_sources_common(c, n, xyz, w, B, normal, hipMemcpyDeviceToDevice); });
}
int fmm_set_weights_dev(fmm_ctx* c, const double* w) {
    return guard(c, [&] {
        require(w != nullptr, FMM_ERR_ARG, "null weights");
        require(!c->bem, FMM_ERR_UNSUPPORTED, "new weights for BEM sources: set the sources again");
        join_default_stream(c);
        if (c->ns) FMM_HIP(hipMemcpyAsync(c->src_w.ptr, w, c->ns * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        if (c->stage >= ST_TREES) {
            int nss = c->S.num_slots;
            launch_charges(c->S.slots.as<int>(), nss, c->src_xyz.as<double>(), c->src_w.as<double>(),
                           c->have_J ? c->src_J.as<double>() : nullptr, c->sxyz.as<double>(), c->sq.as<double>(),
                           c->sJ.as<double>(), c->stream);
            c->stage = ST_TREES;
        }
    });
}

int fmm_build_trees(fmm_ctx* c) { return guard(c, [&] { stage_trees(c); }); }
int fmm_leaf_expansions(fmm_ctx* c) { return guard(c, [&] { stage_leaves(c); }); }
int fmm_upward(fmm_ctx* c) {
    return guard(c, [&] {
        if (c->stage == ST_TREES) stage_leaves(c);   // calclMultipoleExpansions runs P2M if needed (:104-107)
        stage_upward(c);
    });
}
int fmm_upward_with(fmm_ctx* c, int algorithm) {
    return guard(c, [&] {
        require(algorithm >= FMM_M2M_NO_TRANSLATION && algorithm <= FMM_M2M_MATRICES, FMM_ERR_ARG,
                "unknown M2M algorithm");
        if (c->stage == ST_TREES) stage_leaves(c);
        stage_upward(c, algorithm);
    });
}
int fmm_downward(fmm_ctx* c) { return guard(c, [&] { stage_downward(c); }); }

int fmm_evaluate_B_dev(fmm_ctx* c, double current, double* B) {
    return guard(c, [&] {
        require(B != nullptr || c->nt == 0, FMM_ERR_ARG, "null output");
        stage_evaluate(c, current, B);
    });
}
int fmm_evaluate_B(fmm_ctx* c, double current, double* B) {
    return guard(c, [&] {
        require(B != nullptr || c->nt == 0, FMM_ERR_ARG, "null output");
        double* d = c->Bout.get<double>(3 * (size_t)std::max(c->nt, 1));
        stage_evaluate(c, current, d);
        if (c->nt) FMM_HIP(hipMemcpyAsync(B, d, 3 * (size_t)c->nt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        FMM_HIP(hipStreamSynchronize(c->stream));
    });
}
int fmm_run_B_dev(fmm_ctx* c, double current, double* B) {
    return guard(c, [&] { run_fused(c, current, B); });
}

int fmm_leaf_order(fmm_ctx* c, int64_t* order) {
    return guard(c, [&] {
        require(c->stage >= ST_TREES, FMM_ERR_STATE, "trees are not built");
        std::vector<int> s(c->T.num_slots);
        if (!s.empty())
            FMM_HIP(hipMemcpyAsync(s.data(), c->T.slots.ptr, s.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        FMM_HIP(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < s.size(); i++) order[i] = s[i];
    });
}

int fmm_adjoint_weights_dev(fmm_ctx* c, double current, const double* dLdB, double* dLdw) {
    return guard(c, [&] {
        require(c->stage >= ST_DOWN, FMM_ERR_STATE, "run the forward pass first");
        require(dLdB && dLdw, FMM_ERR_ARG, "null adjoint arrays");
        require(c->cfg.frame == FMM_FRAME_SHARED, FMM_ERR_UNSUPPORTED, "the adjoint needs the shared frame");
        require(c->cfg.near_field != FMM_NEAR_TREECODE, FMM_ERR_UNSUPPORTED,
                "the adjoint has no transposed treecode near field");
        require(!c->bem, FMM_ERR_UNSUPPORTED, "the adjoint is for current weights, not BEM sources");
        join_default_stream(c);
        AdjointArgs a;
        a.prec = c->pass_prec();
        if (c->prec == Precision::F16X3 && c->down_ops.nops > c->down_l2l)
            // the transposed passes apply the down operators in FP32: the M2L ones too
            build_tables(c, c->down_ops, c->down_specs, c->stream);
        a.p = c->p;
        a.Lp = c->Lp;
        a.S = &c->S;
        a.T = &c->T;
        a.sigS = c->sigS;
        a.sigT = c->sigT;
        a.src_leaves = c->src_leaves.as<int>();
        a.n_src_leaves = (int)c->h_src_leaves.size();
        a.tgt_leaves = c->tgt_leaves.as<int>();
        a.n_tgt_leaves = (int)c->h_tgt_leaves.size();
        a.sxyz = c->sxyz.as<double>();
        a.sJ = c->sJ.as<double>();
        a.txyz = c->txyz.as<double>();
        a.lists = &c->lists;
        a.near_field = c->cfg.near_field;
        a.up_ops = ops_ptr(c, c->up_ops);
        a.up_ops_lo = c->up_ops.lo.ptr;
        a.down_ops = ops_ptr(c, c->down_ops);
        a.down_ops_lo = c->down_ops.lo.ptr;
        // M2M^T needs (child, parent) pairs per orientation (the forward pass is parent-major)
        const TreeDev& S = c->S;
        Staging& sg = c->adj_stage;
        sg.begin();
        c->m2m_adj_pass.assign(MAX_LEVELS, Pass());
        for (int k = S.num_levels - 1; k >= 1; k--) {
            std::vector<int> src[8], dst[8];
            for (int n = S.level_begin[k - 1]; n < S.level_begin[k]; n++) {
                int fc = S.h_first_child[n];
                if (fc < 0) continue;
                for (int o = 0; o < 8; o++)
                    if (nonempty_host(S, fc + o)) { src[o].push_back(fc + o); dst[o].push_back(n); }
            }
            std::vector<int> all_s, all_d, cb(1, 0), cop;
            for (int o = 0; o < 8; o++) {
                all_s.insert(all_s.end(), src[o].begin(), src[o].end());
                all_d.insert(all_d.end(), dst[o].begin(), dst[o].end());
                cb.push_back((int)all_s.size());
                cop.push_back((k - 1) * 8 + o);
            }
            std::vector<GemmTile> tiles;
            plan_tiles(cb, cop, tiles);
            stage_pass(sg, c->m2m_adj_pass[k], all_s, all_d, tiles);
        }
        // M2L^T runs on class_gemm (tiles of at most 64 columns); the forward M2L tiles may be
        // wider (F16X3 plans 128), so the transposed pass gets its own 64-column tiles
        size_t o_m2l_tiles = 0;
        int n_m2l_tiles = 0;
        if (c->m2l_pass.ntiles) {
            const int nc = c->lists.num_classes;
            std::vector<int> cop(nc);
            for (int i = 0; i < nc; i++) cop[i] = c->down_l2l + i;
            std::vector<GemmTile> tiles;
            plan_tiles(c->lists.h_class_begin, cop, tiles, 64);
            n_m2l_tiles = (int)tiles.size();
            o_m2l_tiles = sg.put(tiles);
        }
        const char* dev = sg.send(c->stream);
        for (int k = S.num_levels - 1; k >= 1; k--) resolve_pass(c->m2m_adj_pass[k], dev);
        for (int k = 0; k < MAX_LEVELS; k++) {
            const Pass* pm = k < (int)c->m2m_adj_pass.size() ? &c->m2m_adj_pass[k] : nullptr;
            const Pass* pl = k < (int)c->l2l_pass.size() ? &c->l2l_pass[k] : nullptr;
            a.m2m_levels.push_back(pm ? PassView{pm->ntiles, pm->tiles, pm->src, pm->dst, pm->width} : PassView{});
            a.l2l_levels.push_back(pl ? PassView{pl->ntiles, pl->tiles, pl->src, pl->dst, pl->width} : PassView{});
        }
        a.m2l = PassView{n_m2l_tiles, (const GemmTile*)(dev + o_m2l_tiles), c->lists.far_t.as<int>(),
                         c->lists.far_s.as<int>(), 64};
        a.n_up_ops = c->up_ops.nops;
        a.n_down_ops = c->down_ops.nops;
        a.scale = current * REF_MU0 / (4.0 * REF_PI);
        a.nt = c->nt;
        a.ns = c->ns;
        run_adjoint(c->adj, a, dLdB, dLdw, c->stream);
    });
}

static int direct_entry(fmm_ctx* c, bool pot, double current, size_t m, const double* xyz, double* out) {
    return guard(c, [&] {
        require(c->stage >= ST_TREES, FMM_ERR_STATE, "trees are not built (sources are gathered with them)");
        require(m < (size_t)INT32_MAX / 3, FMM_ERR_LIMIT, "too many targets");
        require(m == 0 || (xyz && out), FMM_ERR_ARG, "null target or output array");
        if (!m) return;
        join_default_stream(c);
        const int nsplit = direct_splits(c->S.num_slots, (int)m);
        double* part = c->direct_part.get<double>((size_t)nsplit * 3 * m);
        direct_sum(c, pot, current, m, xyz, out, part, nsplit);
        FMM_HIP(hipStreamSynchronize(c->stream));
    });
}
int fmm_direct_B_dev(fmm_ctx* c, double current, size_t m, const double* xyz, double* B) {
    return direct_entry(c, false, current, m, xyz, B);
}
int fmm_direct_A_dev(fmm_ctx* c, double current, size_t m, const double* xyz, double* A) {
    return direct_entry(c, true, current, m, xyz, A);
}
// fmm_direct_B_dev without allocating: the partial sums go to the device scratch array (as many
// source ranges as fit), nothing is allocated or freed
int fmm_direct_B_noalloc_dev(fmm_ctx* c, double current, size_t m, const double* xyz, double* B) {
    return guard(c, [&] {
        require(c->stage >= ST_TREES, FMM_ERR_STATE, "trees are not built (sources are gathered with them)");
        require(m <= FMM_NOALLOC_MAX_TARGETS, FMM_ERR_LIMIT,
                "fmm_direct_B_noalloc_dev: more targets than FMM_NOALLOC_MAX_TARGETS (65536)");
        require(m == 0 || (xyz && B), FMM_ERR_ARG, "null target or output array");
        require(c->scratch != nullptr, FMM_ERR_STATE, "scratch array not set up");
        if (!m) return;
        join_default_stream(c);
        const int cap = (int)(FMM_NOALLOC_SCRATCH_DOUBLES / (3 * m));
        const int nsplit = std::max(1, std::min(direct_splits(c->S.num_slots, (int)m), cap));
        direct_sum(c, false, current, m, xyz, B, c->scratch, nsplit);
        FMM_HIP(hipStreamSynchronize(c->stream));
    });
}

int fmm_evaluate_A_dev(fmm_ctx* c, double current, double* A) {
    return guard(c, [&] {
        require(A != nullptr || c->nt == 0, FMM_ERR_ARG, "null output");
        stage_evaluate_A(c, current, A);
    });
}
int fmm_evaluate_A(fmm_ctx* c, double current, double* A) {
    return guard(c, [&] {
        require(A != nullptr || c->nt == 0, FMM_ERR_ARG, "null output");
        double* d = c->Aout.get<double>(3 * (size_t)std::max(c->nt, 1));
        stage_evaluate_A(c, current, d);
        if (c->nt) FMM_HIP(hipMemcpyAsync(A, d, 3 * (size_t)c->nt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        FMM_HIP(hipStreamSynchronize(c->stream));
    });
}

int fmm_evaluate_bem(fmm_ctx* c, double* out) {
    return guard(c, [&] {
        require(c->bem, FMM_ERR_STATE, "the sources are not BEM sources (fmm_set_bem_sources)");
        require(out != nullptr || c->nt == 0, FMM_ERR_ARG, "null output");
        double* d = c->Aout.get<double>(3 * (size_t)std::max(c->nt, 1));
        stage_evaluate_A(c, 1.0, d);
        if (c->nt) FMM_HIP(hipMemcpyAsync(out, d, 3 * (size_t)c->nt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        FMM_HIP(hipStreamSynchronize(c->stream));
    });
}
int fmm_evaluate_bem_dev(fmm_ctx* c, double* out) {
    return guard(c, [&] {
        require(c->bem, FMM_ERR_STATE, "the sources are not BEM sources (fmm_set_bem_sources)");
        require(out != nullptr || c->nt == 0, FMM_ERR_ARG, "null output");
        stage_evaluate_A(c, 1.0, out);
    });
}

int fmm_treecode_dev(fmm_ctx* c, double current, size_t m, const double* xyz, double* B, double* A) {
    return guard(c, [&] {
        join_default_stream(c);
        stage_treecode(c, current, m, xyz, B, A);
        FMM_HIP(hipStreamSynchronize(c->stream));
    });
}
int fmm_treecode(fmm_ctx* c, double current, size_t m, const double* xyz, double* B, double* A) {
    return guard(c, [&] {
        require(m == 0 || xyz, FMM_ERR_ARG, "null point array");
        require(m < (size_t)INT32_MAX / 3, FMM_ERR_LIMIT, "too many points");
        const size_t n3 = 3 * std::max<size_t>(m, 1);
        double* dx = c->pts.get<double>(n3);
        double* dB = B ? c->tc_B.get<double>(n3) : nullptr;
        double* dA = A ? c->tc_A.get<double>(n3) : nullptr;
        if (m) FMM_HIP(hipMemcpyAsync(dx, xyz, 3 * m * sizeof(double), hipMemcpyHostToDevice, c->stream));
        stage_treecode(c, current, m, dx, dB, dA);
        if (m && B) FMM_HIP(hipMemcpyAsync(B, dB, 3 * m * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        if (m && A) FMM_HIP(hipMemcpyAsync(A, dA, 3 * m * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        FMM_HIP(hipStreamSynchronize(c->stream));
    });
}

int fmm_get_stats(fmm_ctx* c, fmm_stats* s) {
    return guard(c, [&] {
        require(s != nullptr, FMM_ERR_ARG, "null stats");
        std::memset(s, 0, sizeof(*s));
        FMM_HIP(hipStreamSynchronize(c->stream));
        s->source_nodes = c->S.num_nodes;
        s->target_nodes = c->T.num_nodes;
        s->source_levels = c->S.num_levels;
        s->target_levels = c->T.num_levels;
        s->source_slots = c->S.num_slots;
        s->target_slots = c->T.num_slots;
        s->dropped_sources = c->S.dropped;
        s->dropped_targets = c->T.dropped;
        s->far_pairs = c->lists.num_far;
        s->close_pairs = c->lists.num_close;
        s->m2l_classes = c->lists.num_classes;
        s->order = c->p;
        s->padded_length = c->Lp;
        s->bfs_steps = c->lists.bfs_steps;
        if (c->stage >= ST_DOWN && c->near_inter < 0 && c->lists.num_close > 0) {
            std::vector<int> ct(c->lists.num_close), cs(c->lists.num_close);
            FMM_HIP(hipMemcpy(ct.data(), c->lists.close_t.ptr, ct.size() * sizeof(int), hipMemcpyDeviceToHost));
            FMM_HIP(hipMemcpy(cs.data(), c->lists.close_s.ptr, cs.size() * sizeof(int), hipMemcpyDeviceToHost));
            fetch_tree_details(c->T);
            fetch_tree_details(c->S);
            long long tot = 0;
            for (size_t i = 0; i < ct.size(); i++)
                tot += (long long)(c->T.h_end[ct[i]] - c->T.h_begin[ct[i]]) * (c->S.h_end[cs[i]] - c->S.h_begin[cs[i]]);
            c->near_inter = tot;
        }
        s->source_leaves = (int)c->h_src_leaves.size();
        s->target_leaves = (int)c->h_tgt_leaves.size();
        s->source_tree_sorted = c->S.sorted_build ? 1 : 0;
        s->target_tree_sorted = c->T.sorted_build ? 1 : 0;
        s->near_interactions = c->near_inter < 0 ? 0 : c->near_inter;
        auto el = [&](int a, int b) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, c->ev[a], c->ev[b]) != hipSuccess) return 0.0;
            return (double)ms;
        };
        if (c->stage >= ST_EVAL) {
            s->t_tree = el(E_START, E_TREE);
            s->t_p2m = el(E_TREE, E_P2M);
            s->t_m2m = el(E_P2M, E_M2M);
            s->t_lists = el(E_LISTS0, E_LISTS);
            s->t_lists_far = el(E_LISTS0, E_FAR);
            s->t_tables = el(E_DOWN0, E_TABLES);
            s->t_m2l = c->cfg.frame == FMM_FRAME_SHARED ? el(E_GEMM0, E_M2L) : el(E_TABLES, E_M2L);
            s->t_l2l = el(E_M2L, E_L2L);
            s->t_l2p = el(E_L2L, E_L2P);
            s->t_p2p = el(E_P2P0, E_P2P);
            s->t_total = el(E_START, E_END);
            s->t_m2l_gemm = c->cfg.frame == FMM_FRAME_SHARED ? el(E_GEMM0, E_M2L_GEMM) : s->t_m2l;
            if (c->cfg.frame == FMM_FRAME_SHARED) {
                s->t_to_m2l_gemm = el(E_START, E_GEMM0);
                s->t_after_m2l_gemm = el(E_M2L_GEMM, E_END);
            }
        }
    });
}

int fmm_tree_size(fmm_ctx* c, int which, int* nn, int* ns) {
    return guard(c, [&] {
        require(c->stage >= ST_TREES, FMM_ERR_STATE, "trees are not built");
        const TreeDev& T = which ? c->T : c->S;
        *nn = T.num_nodes;
        *ns = T.num_slots;
    });
}

int fmm_tree_export(fmm_ctx* c, int which, int* level, int* parent, int* first_child, int* begin, int* end, int* npts,
                    double* center, int* slots) {
    return guard(c, [&] {
        require(c->stage >= ST_TREES, FMM_ERR_STATE, "trees are not built");
        TreeDev& T = which ? c->T : c->S;
        fetch_tree_details(T);
        int n = T.num_nodes;
        std::memcpy(level, T.h_level.data(), n * sizeof(int));
        std::memcpy(parent, T.h_parent.data(), n * sizeof(int));
        std::memcpy(first_child, T.h_first_child.data(), n * sizeof(int));
        std::memcpy(begin, T.h_begin.data(), n * sizeof(int));
        std::memcpy(end, T.h_end.data(), n * sizeof(int));
        std::memcpy(npts, T.h_npts.data(), n * sizeof(int));
        std::memcpy(center, T.h_center.data(), 3 * (size_t)n * sizeof(double));
        if (T.num_slots)
            FMM_HIP(hipMemcpy(slots, T.slots.ptr, T.num_slots * sizeof(int), hipMemcpyDeviceToHost));
    });
}

int fmm_lists_size(fmm_ctx* c, long long* nfar, long long* nclose, long long* ncls) {
    return guard(c, [&] {
        require(c->lists_ready, FMM_ERR_STATE, "lists are not built");
        *nfar = c->lists.num_far;
        *nclose = c->lists.num_close;
        *ncls = c->lists.num_classes;
    });
}

int fmm_lists_export(fmm_ctx* c, int* ft, int* fs, int* ct, int* cs) {
    return guard(c, [&] {
        require(c->lists_ready, FMM_ERR_STATE, "lists are not built");
        FMM_HIP(hipStreamSynchronize(c->stream));
        const Lists& L = c->lists;
        if (L.num_far) {
            FMM_HIP(hipMemcpy(ft, L.far_t.ptr, L.num_far * sizeof(int), hipMemcpyDeviceToHost));
            FMM_HIP(hipMemcpy(fs, L.far_s.ptr, L.num_far * sizeof(int), hipMemcpyDeviceToHost));
        }
        if (L.num_close) {
            FMM_HIP(hipMemcpy(ct, L.close_t.ptr, L.num_close * sizeof(int), hipMemcpyDeviceToHost));
            FMM_HIP(hipMemcpy(cs, L.close_s.ptr, L.num_close * sizeof(int), hipMemcpyDeviceToHost));
        }
    });
}

int fmm_export_B_parts(fmm_ctx* c, double* Bfar, double* Bnear) {
    return guard(c, [&] {
        require(c->stage >= ST_EVAL, FMM_ERR_STATE, "no evaluation yet");
        FMM_HIP(hipStreamSynchronize(c->stream));
        size_t n = 3 * (size_t)c->T.num_slots;
        double sc = c->current * REF_MU0 / (4.0 * REF_PI);
        if (n) {
            FMM_HIP(hipMemcpy(Bfar, c->Bfar.ptr, n * sizeof(double), hipMemcpyDeviceToHost));
            FMM_HIP(hipMemcpy(Bnear, c->Bnear.ptr, n * sizeof(double), hipMemcpyDeviceToHost));
        }
        for (size_t i = 0; i < n; i++) { Bfar[i] *= sc; Bnear[i] *= sc; }
    });
}

int fmm_export_expansion(fmm_ctx* c, int which, int node, double* out) {
    return guard(c, [&] {
        require(which ? c->stage >= ST_DOWN : c->stage >= ST_LEAVES, FMM_ERR_STATE, "expansion not computed");
        TreeDev& T = which ? c->T : c->S;
        require(node >= 0 && node < T.num_nodes, FMM_ERR_ARG, "node out of range");
        FMM_HIP(hipStreamSynchronize(c->stream));
        fetch_tree_details(T);
        const DevBuf& E = which ? c->Lx : c->M;
        std::vector<double> v(3 * (size_t)c->Lp);
        if (c->f64()) {
            FMM_HIP(hipMemcpy(v.data(), (const double*)E.ptr + (size_t)node * 3 * c->Lp, v.size() * sizeof(double),
                              hipMemcpyDeviceToHost));
        } else {
            std::vector<float> f(v.size());
            FMM_HIP(hipMemcpy(f.data(), (const float*)E.ptr + (size_t)node * 3 * c->Lp, f.size() * sizeof(float),
                              hipMemcpyDeviceToHost));
            for (size_t i = 0; i < f.size(); i++) v[i] = f[i];
        }
        double sg = T.scale[T.h_level[node]];
        for (int comp = 0; comp < 3; comp++)
            for (int l = 0; l <= c->p; l++)
                for (int m = -l; m <= l; m++) {
                    int k = l * l + l + m;
                    double f = which ? std::pow(sg, -(l + 1)) : std::pow(sg, l);
                    out[comp * c->L + k] = v[comp * c->Lp + k] * f;
                }
    });
}

int fmm_make_operator(int kind, int order, const double* t, double sig_in, double sig_out, int m2l_full, double* out) {
    if (kind < 0 || kind > 2 || order < 1 || order > MAX_ORDER || !t || !out) return FMM_ERR_ARG;
    try {
        int L = nlen(order), Lp = pad_to(L, 64);
        OpSpec s;
        for (int a = 0; a < 3; a++) s.t[a] = t[a];
        s.sig_in = sig_in;
        s.sig_out = sig_out;
        s.kind = kind;
        s.m2l_full = m2l_full;
        DevBuf d, o;
        OpSpec* ds = d.get<OpSpec>(1);
        FMM_HIP(hipMemcpy(ds, &s, sizeof(s), hipMemcpyHostToDevice));
        double* dout = o.get<double>((size_t)Lp * Lp);
        make_operators(ds, 1, order, Lp, 0, dout, nullptr, 0);
        std::vector<double> h((size_t)Lp * Lp);
        FMM_HIP(hipMemcpy(h.data(), dout, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < L; i++)
            for (int j = 0; j < L; j++) out[(size_t)i * L + j] = h[(size_t)i * Lp + j];
        return FMM_OK;
    } catch (const Error& e) {
        return e.code;
    } catch (...) {
        return FMM_ERR_HIP;
    }
}

}  // extern "C"

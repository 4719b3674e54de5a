// Class-batched translation GEMMs on MFMA (gfx950).
//
// A "class" is one translation operator T (Lp x Lp, row = output coefficient) shared
// by a list of (source node, destination node) pairs: an M2L offset class, or one M2M /
// L2L orientation of a level.  Column j of a class is component j % 3 of pair j / 3:
// B[:, j] = src_exp[src][comp][:], and T B[:, j] is added to dst_exp[dst][comp][:].  One
// launch covers every 64-column tile of every class (tiles never straddle classes).
// Expansions are node-major [node][3][Lp] so a column is one contiguous vector.
//
// Each workgroup (4 waves) computes the whole Lp-row panel of its 64 columns, so every
// gathered column is read from HBM once; wave w owns the row blocks w, w + 4, ... (RB of them,
// Lp = 64 RB) times the 4 column blocks, as MFMA blocks of 16 x 16.  K is staged through double-buffered LDS with the
// next step's global loads issued before the current step's MFMAs (one barrier per step).
//
// Precisions:
//   F64    : v_mfma_f64_16x16x4f64 (C register r holds row 4r + lane/16);
//   F32    : v_mfma_f32_16x16x4f32 (C register r holds row 4(lane/16) + r), FP32 accumulate;
//   BF16X3 : v_mfma_f32_16x16x32_bf16 on A = Ah + Al (tables split offline from FP64) and
//            B = Bh + Bl (split on load): C += Al Bh + Ah Bl + Ah Bh, FP32 accumulate.
// TRANS applies T^T (the adjoint passes) by reading the operator column-wise.
#include "fmm_gemm.hpp"
#include "fmm_prims.hpp"

namespace fmm {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int BN = 64;

// hardware float atomics (global_atomic_add_f32/f64) on coarse-grained device memory
__device__ __forceinline__ void add_out(float* p, float v, int epi) {
    if (epi == EPI_ATOMIC) unsafeAtomicAdd(p, v);
    else *p += v;
}
__device__ __forceinline__ void add_out(double* p, double v, int epi) {
    if (epi == EPI_ATOMIC) unsafeAtomicAdd(p, v);
    else *p += v;
}

template <class T> struct Vec4;
template <> struct Vec4<float> {
    typedef float4 t;
    static __device__ __forceinline__ float4 load(const float* p) { return *(const float4*)p; }
    static __device__ __forceinline__ float get(const float4& v, int i) { return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w; }
    static __device__ __forceinline__ float4 zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
};
template <> struct Vec4<double> {
    struct t { double2 a, b; };
    static __device__ __forceinline__ t load(const double* p) { return t{*(const double2*)p, *(const double2*)(p + 2)}; }
    static __device__ __forceinline__ double get(const t& v, int i) { return i == 0 ? v.a.x : i == 1 ? v.a.y : i == 2 ? v.b.x : v.b.y; }
    static __device__ __forceinline__ t zero() { return t{make_double2(0.0, 0.0), make_double2(0.0, 0.0)}; }
};

// ---------------------------------------------------------------- FP64 / FP32
template <class T, int RB, bool TRANS, bool MULTI>
__global__ __launch_bounds__(256) void k_gemm(const T* __restrict__ ops, const GemmTile* __restrict__ tiles,
                                              const int* __restrict__ pair_src, const int* __restrict__ pair_dst,
                                              const T* __restrict__ src_exp, T* __restrict__ dst_exp, int epi,
                                              int nseg, int tri) {
    // LDS row strides: four K-rows are read by the 16-lane groups of a wave at once; strides of
    // 16 (mod 64) words put the four rows on disjoint bank ranges (FP64: 8 doubles = 16 words)
    constexpr int Lp = 64 * RB, BK = 16, PAD = sizeof(T) == 8 ? 8 : 16, SA = Lp + PAD, SB = BN + PAD, NK = Lp / BK;
    const int NKT = MULTI ? nseg * NK : NK;   // single-segment launches keep a compile-time trip count
    typedef Vec4<T> V;
    typedef typename V::t v4;
    typedef typename std::conditional<sizeof(T) == 8, f64x4, f32x4>::type acc_t;
    __shared__ T As[2][BK][SA];
    __shared__ T Bs[2][BK][SB];
    const GemmTile tile = tiles[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // B loader: one column and 4 consecutive k per thread; segment g reads node pair_src + g
    const int bc = tid >> 2, bq = (tid & 3) * 4;
    const T* bptr = nullptr;
    if (bc < tile.ncol) {
        int col = tile.col0 + bc;
        int pair = col / 3, comp = col - 3 * pair;
        bptr = src_exp + ((size_t)pair_src[pair] * 3 + comp) * Lp;
    }
    // A loader: RB quads of 4 per thread.  Non-transposed: (row, 4 consecutive k);
    // transposed: (4 consecutive rows, one k) of T^T = columns of T.  Segment g applies
    // operator tile.op + g.
    v4 ra[RB], rb;
    auto gload = [&](int kt) {
        const int seg = kt / NK, k0 = (kt - seg * NK) * BK;
        const T* A = ops + (size_t)(tile.op + seg) * Lp * Lp;
#pragma unroll
        for (int j = 0; j < RB; j++) {
            int qi = tid + 256 * j;
            if (!TRANS) {
                int row = qi >> 2, kq = (qi & 3) * 4;
                ra[j] = V::load(A + (size_t)row * Lp + k0 + kq);
            } else {
                int k = qi / (Lp / 4), rq = (qi % (Lp / 4)) * 4;
                ra[j] = V::load(A + (size_t)(k0 + k) * Lp + rq);
            }
        }
        if (bptr) rb = V::load(bptr + (size_t)seg * 3 * Lp + k0 + bq);
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int j = 0; j < RB; j++) {
            int qi = tid + 256 * j;
            if (!TRANS) {
                int row = qi >> 2, kq = (qi & 3) * 4;
                for (int i = 0; i < 4; i++) As[buf][kq + i][row] = V::get(ra[j], i);
            } else {
                int k = qi / (Lp / 4), rq = (qi % (Lp / 4)) * 4;
                for (int i = 0; i < 4; i++) As[buf][k][rq + i] = V::get(ra[j], i);
            }
        }
        for (int i = 0; i < 4; i++) Bs[buf][bq + i][bc] = bptr ? V::get(rb, i) : T(0);
    };
    acc_t acc[RB][4];
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = acc_t{0, 0, 0, 0};
    gload(0);
    lstore(0);
    __syncthreads();
    // wave w owns the row blocks w, w + 4, ..., so under a degree-triangular operator every wave
    // skips about as many blocks as the others; per block its lowest and highest degree
    int dlo[RB], dhi[RB];
#pragma unroll
    for (int a = 0; a < RB; a++) {
        const int r0 = 16 * (wave + 4 * a);
        dlo[a] = (int)sqrtf((float)r0);
        dhi[a] = (int)sqrtf((float)(r0 + 15));
    }
    for (int kt = 0; kt < NKT; kt++) {
        const int cur = kt & 1;
        if (kt + 1 < NKT) gload(kt + 1);
        const int k0 = (kt % NK) * BK, klo = (int)sqrtf((float)k0), khi = (int)sqrtf((float)(k0 + BK - 1));
        bool live[RB];
#pragma unroll
        for (int a = 0; a < RB; a++)
            live[a] = tri == TRI_NONE || (tri == TRI_LOWER ? klo <= dhi[a] : khi >= dlo[a]);
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            T af[RB], bf[4];
            const int kk = ks + (lane >> 4);
#pragma unroll
            for (int b = 0; b < 4; b++) bf[b] = Bs[cur][kk][b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < RB; a++) {
                if (!live[a]) continue;   // uniform over the wave
                af[a] = As[cur][kk][16 * (wave + 4 * a) + (lane & 15)];
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    if constexpr (sizeof(T) == 8)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
                    else
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], bf[b], acc[a][b], 0, 0, 0);
                }
            }
        }
        if (kt + 1 < NKT) lstore(cur ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int b = 0; b < 4; b++) {
        int cl = b * 16 + (lane & 15);
        if (cl >= tile.ncol) continue;
        int col = tile.col0 + cl;
        if (epi == EPI_STORE) {
            T* out = dst_exp + (size_t)col * Lp;
#pragma unroll
            for (int a = 0; a < RB; a++) {
                const int r0 = 16 * (wave + 4 * a);
                if constexpr (sizeof(T) == 8) {
#pragma unroll
                    for (int r = 0; r < 4; r++) out[r0 + 4 * r + (lane >> 4)] = acc[a][b][r];
                } else {
                    *(float4*)(out + r0 + 4 * (lane >> 4)) =
                        make_float4(acc[a][b][0], acc[a][b][1], acc[a][b][2], acc[a][b][3]);
                }
            }
            continue;
        }
        int pair = col / 3, comp = col - 3 * pair;
        T* dst = dst_exp + ((size_t)pair_dst[pair] * 3 + comp) * Lp;
        if constexpr (sizeof(T) == 4) {
            if (epi == EPI_ADD) {   // a lane's four rows are consecutive: one 16-byte read-modify-write
#pragma unroll
                for (int a = 0; a < RB; a++) {
                    float4* q = (float4*)(dst + 16 * (wave + 4 * a) + 4 * (lane >> 4));
                    float4 v = *q;
                    v.x += acc[a][b][0];
                    v.y += acc[a][b][1];
                    v.z += acc[a][b][2];
                    v.w += acc[a][b][3];
                    *q = v;
                }
                continue;
            }
        }
#pragma unroll
        for (int a = 0; a < RB; a++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int rr = (sizeof(T) == 8) ? (4 * r + (lane >> 4)) : (4 * (lane >> 4) + r);
                add_out(dst + 16 * (wave + 4 * a) + rr, (T)acc[a][b][r], epi);
            }
    }
}

// ---------------------------------------------------------------- BF16x3
template <int RB, bool TRANS, bool MULTI>
__global__ __launch_bounds__(256) void k_gemm_bf16x3(const __bf16* __restrict__ ops_hi, const __bf16* __restrict__ ops_lo,
                                                     const GemmTile* __restrict__ tiles,
                                                     const int* __restrict__ pair_src, const int* __restrict__ pair_dst,
                                                     const float* __restrict__ src_exp, float* __restrict__ dst_exp,
                                                     int epi, int nseg) {
    constexpr int Lp = 64 * RB, BK = 32, S = BK + 8, NK = Lp / BK;
    const int NKT = MULTI ? nseg * NK : NK;
    // k-contiguous rows: each lane reads its 8 consecutive k as one 16-byte vector
    __shared__ __attribute__((aligned(16))) __bf16 Ah[2][Lp][S];
    __shared__ __attribute__((aligned(16))) __bf16 Al[2][Lp][S];
    __shared__ __attribute__((aligned(16))) __bf16 Bh[2][BN][S];
    __shared__ __attribute__((aligned(16))) __bf16 Bl[2][BN][S];
    const GemmTile tile = tiles[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bc = tid >> 2, bq = (tid & 3) * 8;
    const float* bptr = nullptr;
    if (bc < tile.ncol) {
        int col = tile.col0 + bc;
        int pair = col / 3, comp = col - 3 * pair;
        bptr = src_exp + ((size_t)pair_src[pair] * 3 + comp) * Lp;
    }
    // A loader: RB units of 8 bf16 per thread and part
    bf16x8 rah[RB], ral[RB];
    float4 rb0 = {0, 0, 0, 0}, rb1 = {0, 0, 0, 0};
    auto gload = [&](int kt) {
        const int seg = kt / NK, k0 = (kt - seg * NK) * BK;
        const size_t opoff = (size_t)(tile.op + seg) * Lp * Lp;
        const __bf16* Ahg = ops_hi + opoff;
        const __bf16* Alg = ops_lo + opoff;
#pragma unroll
        for (int j = 0; j < RB; j++) {
            int qi = tid + 256 * j;
            if (!TRANS) {
                int row = qi >> 2, kq = (qi & 3) * 8;
                rah[j] = *(const bf16x8*)(Ahg + (size_t)row * Lp + k0 + kq);
                ral[j] = *(const bf16x8*)(Alg + (size_t)row * Lp + k0 + kq);
            } else {
                int k = qi / (Lp / 8), rq = (qi % (Lp / 8)) * 8;
                rah[j] = *(const bf16x8*)(Ahg + (size_t)(k0 + k) * Lp + rq);
                ral[j] = *(const bf16x8*)(Alg + (size_t)(k0 + k) * Lp + rq);
            }
        }
        if (bptr) {
            const float* bp = bptr + (size_t)seg * 3 * Lp + k0 + bq;
            rb0 = *(const float4*)bp;
            rb1 = *(const float4*)(bp + 4);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int j = 0; j < RB; j++) {
            int qi = tid + 256 * j;
            if (!TRANS) {
                int row = qi >> 2, kq = (qi & 3) * 8;
                *(bf16x8*)&Ah[buf][row][kq] = rah[j];
                *(bf16x8*)&Al[buf][row][kq] = ral[j];
            } else {
                int k = qi / (Lp / 8), rq = (qi % (Lp / 8)) * 8;
                for (int i = 0; i < 8; i++) {
                    Ah[buf][rq + i][k] = rah[j][i];
                    Al[buf][rq + i][k] = ral[j][i];
                }
            }
        }
        float v[8] = {rb0.x, rb0.y, rb0.z, rb0.w, rb1.x, rb1.y, rb1.z, rb1.w};
        bf16x8 h, l;
        for (int i = 0; i < 8; i++) {
            float x = bptr ? v[i] : 0.0f;
            __bf16 hh = (__bf16)x;
            h[i] = hh;
            l[i] = (__bf16)(x - (float)hh);
        }
        *(bf16x8*)&Bh[buf][bc][bq] = h;
        *(bf16x8*)&Bl[buf][bc][bq] = l;
    };
    f32x4 acc[RB][4];
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = f32x4{0, 0, 0, 0};
    gload(0);
    lstore(0);
    __syncthreads();
    const int wrow = wave * 16 * RB;
    const int kk = 8 * (lane >> 4);
    for (int kt = 0; kt < NKT; kt++) {
        const int cur = kt & 1;
        if (kt + 1 < NKT) gload(kt + 1);
        bf16x8 bh[4], bl[4];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            bh[b] = *(const bf16x8*)&Bh[cur][b * 16 + (lane & 15)][kk];
            bl[b] = *(const bf16x8*)&Bl[cur][b * 16 + (lane & 15)][kk];
        }
#pragma unroll
        for (int a = 0; a < RB; a++) {
            int r = wrow + a * 16 + (lane & 15);
            bf16x8 ah = *(const bf16x8*)&Ah[cur][r][kk];
            bf16x8 al = *(const bf16x8*)&Al[cur][r][kk];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[b], acc[a][b], 0, 0, 0);
            }
        }
        if (kt + 1 < NKT) lstore(cur ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int b = 0; b < 4; b++) {
        int cl = b * 16 + (lane & 15);
        if (cl >= tile.ncol) continue;
        int col = tile.col0 + cl;
        if (epi == EPI_STORE) {
            float* out = dst_exp + (size_t)col * Lp + wrow;
#pragma unroll
            for (int a = 0; a < RB; a++)
                *(float4*)(out + a * 16 + 4 * (lane >> 4)) =
                    make_float4(acc[a][b][0], acc[a][b][1], acc[a][b][2], acc[a][b][3]);
            continue;
        }
        int pair = col / 3, comp = col - 3 * pair;
        float* dst = dst_exp + ((size_t)pair_dst[pair] * 3 + comp) * Lp + wrow;
        if (epi == EPI_ADD) {   // a lane's four rows are consecutive: one 16-byte read-modify-write
#pragma unroll
            for (int a = 0; a < RB; a++) {
                float4* q = (float4*)(dst + a * 16 + 4 * (lane >> 4));
                float4 v = *q;
                v.x += acc[a][b][0];
                v.y += acc[a][b][1];
                v.z += acc[a][b][2];
                v.w += acc[a][b][3];
                *q = v;
            }
            continue;
        }
#pragma unroll
        for (int a = 0; a < RB; a++)
#pragma unroll
            for (int r = 0; r < 4; r++) add_out(dst + a * 16 + 4 * (lane >> 4) + r, acc[a][b][r], epi);
    }
}

// ---------------------------------------------------------------- M2L, rows padded to 16
// The M2L launch with rows and K padded only to R = 16 NB (>= L, <= Lp) instead of Lp: wave w
// owns the 16 columns [16w, 16w + 16) of the 64-column tile and all NB row blocks, so an odd
// block count leaves no wave idle.  Each column is written to the per-pair slab with row
// stride R (rows >= L are zero: the operators are zero there).
constexpr int lds_stride(int rows, int words_per_elem) {
    // rows + pad with the stride = 16 (mod 64) 4-byte words: the four K-rows read at once by the
    // 16-lane groups of a wave fall on disjoint LDS banks
    return words_per_elem == 1 ? rows + ((16 - rows % 64) + 64) % 64 : rows + ((8 - rows % 32) + 32) % 32;
}

template <class T, int NB>
__global__ __launch_bounds__(256) void k_gemm_m2l(const T* __restrict__ ops, const GemmTile* __restrict__ tiles,
                                                  const int* __restrict__ pair_src, const T* __restrict__ src_exp,
                                                  T* __restrict__ slab, int Lp) {
    constexpr int R = 16 * NB, BK = 16, NK = NB;
    // k-contiguous LDS rows: the 16-lane group g of a wave takes k = 4g .. 4g+3 of a K step, one
    // per MFMA (the same order for A and B, so the sum is unchanged), read as one 4-vector
    constexpr int SK = BK + 16 / (int)sizeof(T);   // row stride in elements (pad 16 bytes)
    constexpr int NQ = (NB * 64 + 255) / 256;      // A quads (4 elements) per thread and K step
    typedef Vec4<T> V;
    typedef typename V::t v4;
    typedef typename std::conditional<sizeof(T) == 8, f64x4, f32x4>::type acc_t;
    __shared__ __attribute__((aligned(16))) T As[2][R][SK];
    __shared__ __attribute__((aligned(16))) T Bs[2][BN][SK];
    // workgroups are dealt round-robin to the 8 XCDs: give each XCD a contiguous eighth of the
    // tiles, so the tiles of one class (one operator) share an L2 (a bijection in any case)
    const int nt = gridDim.x, xcd = blockIdx.x & 7, q = nt >> 3, rem = nt & 7;
    const int t = xcd * q + min(xcd, rem) + (blockIdx.x >> 3);
    const GemmTile tile = tiles[t];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = lane >> 4, li = lane & 15;
    const T* A = ops + (size_t)tile.op * Lp * Lp;
    const int bc = tid >> 2, bq = (tid & 3) * 4;
    const T* bptr = nullptr;
    if (bc < tile.ncol) {
        int col = tile.col0 + bc;
        int pair = col / 3, comp = col - 3 * pair;
        bptr = src_exp + ((size_t)pair_src[pair] * 3 + comp) * Lp;
    }
    v4 ra[NQ], rb;
    auto gload = [&](int k0) {
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            int qi = tid + 256 * j;
            if (qi < NB * 64) ra[j] = V::load(A + (size_t)(qi >> 2) * Lp + k0 + (qi & 3) * 4);
        }
        if (bptr) rb = V::load(bptr + k0 + bq);
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            int qi = tid + 256 * j;
            if (qi < NB * 64) *(v4*)&As[buf][qi >> 2][(qi & 3) * 4] = ra[j];
        }
        if (bptr) *(v4*)&Bs[buf][bc][bq] = rb;
        else *(v4*)&Bs[buf][bc][bq] = V::zero();
    };
    acc_t acc[NB];
#pragma unroll
    for (int a = 0; a < NB; a++) acc[a] = acc_t{0, 0, 0, 0};
    gload(0);
    lstore(0);
    __syncthreads();
    for (int kt = 0; kt < NK; kt++) {
        const int cur = kt & 1;
        if (kt + 1 < NK) gload((kt + 1) * BK);
        const v4 bf = *(const v4*)&Bs[cur][wave * 16 + li][4 * g];
        v4 af[NB];
#pragma unroll
        for (int a = 0; a < NB; a++) af[a] = *(const v4*)&As[cur][a * 16 + li][4 * g];
        // consecutive MFMAs update different accumulators (no back-to-back dependence)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int a = 0; a < NB; a++) {
                if constexpr (sizeof(T) == 8)
                    acc[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(V::get(af[a], j), V::get(bf, j), acc[a], 0, 0, 0);
                else
                    acc[a] = __builtin_amdgcn_mfma_f32_16x16x4f32(V::get(af[a], j), V::get(bf, j), acc[a], 0, 0, 0);
            }
        if (kt + 1 < NK) lstore(cur ^ 1);
        __syncthreads();
    }
    // the slab is written once and read once (by the reduction): nontemporal stores keep the
    // operators and gathered multipoles resident in L2 (measured -9 % launch time on the 1M
    // points workload against plain stores)
    const int cl = wave * 16 + li;
    if (cl >= tile.ncol) return;
    T* out = slab + (size_t)(tile.col0 + cl) * R;
#pragma unroll
    for (int a = 0; a < NB; a++) {
        if constexpr (sizeof(T) == 8) {
#pragma unroll
            for (int r = 0; r < 4; r++) __builtin_nontemporal_store(acc[a][r], out + a * 16 + 4 * r + g);
        } else {
            __builtin_nontemporal_store(acc[a], (f32x4*)(out + a * 16 + 4 * g));
        }
    }
}

// ---------------------------------------------------------------- M2L, FP16 split
// The M2L of the F16X3 precision: every product of the FP32 GEMM is replaced by three FP16
// products on v_mfma_f32_16x16x32_f16 (FP32 accumulate), A = Ah + Al and B = Bh + Bl with
// hi and lo in FP16 (22 significant bits together), C = Al Bh + Ah Bl + Ah Bh.  FP16 has a
// narrow range, so both sides are equilibrated by powers of 2 (exact):
//   A_eq = diag(rs) T diag(1/dm)   (operator tables: row scale rs per operator, dm per input
//                                   coefficient; the tables hold A_eq split, rinv = 1 / rs)
//   B_eq = diag(dm) M diag(s)      (multipoles: dm per coefficient, s per column, split once per
//                                   evaluation by k_split_columns into FP16 hi and lo parts)
// so T M = diag(rinv) (A_eq B_eq) diag(1/s), applied in the epilogue.  Rows R = 16 NB >= L,
// K padded to Kp = 32 NKS >= L (zero operator columns and multipole rows beyond L).
// Layout and schedule follow k_gemm_m2l: wave w owns the tile's columns [16w, 16w + 16) and all
// NB row blocks; one LDS buffer refilled from registers whose loads overlap the MFMAs.
typedef _Float16 half8 __attribute__((ext_vector_type(8)));

template <int NB, int NKS, int NC>
__global__ __launch_bounds__(256) void k_gemm_m2l_f16x3(const _Float16* __restrict__ ops_hi,
                                                        const _Float16* __restrict__ ops_lo,
                                                        const float* __restrict__ ops_rinv,
                                                        const GemmTile* __restrict__ tiles,
                                                        const int* __restrict__ pair_src,
                                                        const _Float16* __restrict__ src_hl,
                                                        const float* __restrict__ cscale, float* __restrict__ slab,
                                                        int op_base) {
    constexpr int R = 16 * NB, Kp = 32 * NKS, BK = 32, S = BK + 8;   // LDS row stride: 80 bytes
    constexpr int NQ = (4 * R + 255) / 256;                          // A: 8-half vectors per thread and part
    constexpr int W = 64 * NC;                                       // tile columns (NC blocks of 16 per wave)
    __shared__ __attribute__((aligned(16))) _Float16 Ah[R][S];
    __shared__ __attribute__((aligned(16))) _Float16 Al[R][S];
    __shared__ __attribute__((aligned(16))) _Float16 Bh[W][S];
    __shared__ __attribute__((aligned(16))) _Float16 Bl[W][S];
    // each XCD works on a contiguous eighth of the tiles (a bijection of the grid)
    const int nt = gridDim.x, xcd = blockIdx.x & 7, q = nt >> 3, rem = nt & 7;
    const int t = xcd * q + min(xcd, rem) + (blockIdx.x >> 3);
    const GemmTile tile = tiles[t];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = lane >> 4, li = lane & 15;
    const int op = tile.op - op_base;   // tile.op numbers the operators of all down passes
    const _Float16* Ahg = ops_hi + (size_t)op * R * Kp;
    const _Float16* Alg = ops_lo + (size_t)op * R * Kp;
    // B staging: columns bc + 64 c, k = bq .. bq + 7 of each K step, from the multipoles split
    // beforehand (split_columns): equilibrated FP16 hi and lo parts, [column][K step][hi | lo]
    const int bc = tid >> 2, bq = (tid & 3) * 8;
    const _Float16* bph[NC];
    const _Float16* bpl[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        bph[c] = bpl[c] = nullptr;
        if (bc + 64 * c < tile.ncol) {
            const int col = tile.col0 + bc + 64 * c, pair = col / 3, comp = col - 3 * pair;
            const size_t sc = (size_t)pair_src[pair] * 3 + comp;
            bph[c] = src_hl + sc * 2 * Kp;
            bpl[c] = bph[c] + 32;
        }
    }
    half8 rah[NQ], ral[NQ];
    half8 rbh[NC], rbl[NC];
    auto gload = [&](int k0) {
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            const int qi = tid + 256 * j;
            if (qi < 4 * R) {
                const size_t o = (size_t)(qi >> 2) * Kp + k0 + (qi & 3) * 8;
                rah[j] = *(const half8*)(Ahg + o);
                ral[j] = *(const half8*)(Alg + o);
            }
        }
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const half8 z = {0, 0, 0, 0, 0, 0, 0, 0};
            rbh[c] = bph[c] ? *(const half8*)(bph[c] + 2 * k0 + bq) : z;
            rbl[c] = bpl[c] ? *(const half8*)(bpl[c] + 2 * k0 + bq) : z;
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            const int qi = tid + 256 * j;
            if (qi < 4 * R) {
                *(half8*)&Ah[qi >> 2][(qi & 3) * 8] = rah[j];
                *(half8*)&Al[qi >> 2][(qi & 3) * 8] = ral[j];
            }
        }
#pragma unroll
        for (int c = 0; c < NC; c++) {
            *(half8*)&Bh[bc + 64 * c][bq] = rbh[c];
            *(half8*)&Bl[bc + 64 * c][bq] = rbl[c];
        }
    };
    f32x4 acc[NB][NC];
#pragma unroll
    for (int a = 0; a < NB; a++)
#pragma unroll
        for (int c = 0; c < NC; c++) acc[a][c] = f32x4{0, 0, 0, 0};
    gload(0);
    for (int kt = 0; kt < NKS; kt++) {
        lstore();
        __syncthreads();
        if (kt + 1 < NKS) gload((kt + 1) * BK);
        const int kk = 8 * g;
        half8 bh[NC], bl[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            bh[c] = *(const half8*)&Bh[wave * 16 * NC + 16 * c + li][kk];
            bl[c] = *(const half8*)&Bl[wave * 16 * NC + 16 * c + li][kk];
        }
#pragma unroll
        for (int a = 0; a < NB; a++) {
            const half8 ah = *(const half8*)&Ah[a * 16 + li][kk];
            const half8 al = *(const half8*)&Al[a * 16 + li][kk];
#pragma unroll
            for (int c = 0; c < NC; c++) {
                acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[c], acc[a][c], 0, 0, 0);
                acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[c], acc[a][c], 0, 0, 0);
                acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[c], acc[a][c], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // epilogue: undo the scales; the slab is written once and read once (nontemporal)
    const float* rv = ops_rinv + (size_t)op * R;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int cl = wave * 16 * NC + 16 * c + li;
        if (cl >= tile.ncol) continue;
        const int col = tile.col0 + cl, pair = col / 3, comp = col - 3 * pair;
        const float sinv = 1.0f / cscale[pair_src[pair] * 3 + comp];
        float* out = slab + (size_t)col * R;
#pragma unroll
        for (int a = 0; a < NB; a++) {
            const float4 r = *(const float4*)(rv + a * 16 + 4 * g);
            const f32x4 v = f32x4{acc[a][c][0] * r.x * sinv, acc[a][c][1] * r.y * sinv, acc[a][c][2] * r.z * sinv,
                                  acc[a][c][3] * r.w * sinv};
            __builtin_nontemporal_store(v, (f32x4*)(out + a * 16 + 4 * g));
        }
    }
}

// The 128-column F16X3 M2L with the waves on a 2 x 2 grid: wave (wr, wc) owns the row blocks
// [wr NB0, wr NB0 + NB0) below NB (NB0 = ceil(NB / 2)) and the tile columns [64 wc, 64 wc + 64),
// so it reads half the operator rows and half the columns from LDS per K step (a wave tile of
// 16 NB0 x 64 instead of 16 NB x 32: about a third less LDS traffic per MFMA).  Staging, numerics
// and epilogue as k_gemm_m2l_f16x3.
template <int NB, int NKS, int BD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_num_vgpr(216))) void k_gemm_m2l_f16x3_q(
                                                          const _Float16* __restrict__ ops_hi,
                                                          const _Float16* __restrict__ ops_lo,
                                                          const float* __restrict__ ops_rinv,
                                                          const GemmTile* __restrict__ tiles,
                                                          const int* __restrict__ pair_src,
                                                          const _Float16* __restrict__ src_hl,
                                                          const float* __restrict__ cscale, float* __restrict__ slab,
                                                          int op_base, int ntiles, int* __restrict__ xq) {
    constexpr int R = 16 * NB, Kp = 32 * NKS, BK = 32, S = BK + 8;   // LDS row stride: 80 bytes
    constexpr int NQ = (4 * R + 255) / 256;                          // A: 8-half vectors per thread and part
    constexpr int W = 128, NB0 = (NB + 1) / 2;
    __shared__ __attribute__((aligned(16))) _Float16 Ah[2][R][S];
    __shared__ __attribute__((aligned(16))) _Float16 Al[2][R][S];
    __shared__ __attribute__((aligned(16))) _Float16 Bh[2][W][S];
    __shared__ __attribute__((aligned(16))) _Float16 Bl[2][W][S];
    // persistent: the grid (a multiple of 8, workgroup b on XCD b mod 8) walks the tiles, each
    // XCD a contiguous eighth of them (one class's operator shared in one L2) taken from that
    // XCD's counter xq[xcd] (zero at launch), so a workgroup that starts late (its CU busy)
    // takes fewer tiles, and then from the other XCDs' eighths; a workgroup starts its next tile
    // while its last stores drain.  The
    // register budget leaves room for one near-field workgroup (k_p2p_queue) per CU.
    const int xcd = blockIdx.x & 7, q = ntiles >> 3, rem = ntiles & 7;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = lane >> 4, li = lane & 15;
    const int wr = wave & 1, wc = wave >> 1, rb0 = wr * NB0;
    __shared__ int next_tile;
    int from = xcd;   // (thread 0) the eighth this workgroup draws from: its XCD's, then the others'
    for (;;) {
        if (tid == 0) {
            // once its own eighth is done a workgroup takes tiles left in the others (in turn), so
            // XCDs that finish early help the rest; each counter hands out each index once
            int t = -1;
            for (int s = 0; s < 8 && t < 0; s++) {
                const int x = (from + s) & 7, cnt = q + (x < rem ? 1 : 0);
                const int i = atomicAdd(xq + x, 1);
                if (i < cnt) {
                    t = x * q + min(x, rem) + i;
                    from = x;
                }
            }
            next_tile = t;
        }
        __syncthreads();
        const int it = next_tile;   // read before the K loop's barriers; rewritten only after them
        if (it < 0) break;
        const GemmTile tile = tiles[it];
        const int op = tile.op - op_base;
        const _Float16* Ahg = ops_hi + (size_t)op * R * Kp;
        const _Float16* Alg = ops_lo + (size_t)op * R * Kp;
        const int bc = tid >> 2, bq = (tid & 3) * 8;
        const _Float16* bph[2];
        const _Float16* bpl[2];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            bph[c] = bpl[c] = nullptr;
            if (bc + 64 * c < tile.ncol) {
                const int col = tile.col0 + bc + 64 * c, pair = col / 3, comp = col - 3 * pair;
                const size_t sc = (size_t)pair_src[pair] * 3 + comp;
                bph[c] = src_hl + sc * 2 * Kp;
                bpl[c] = bph[c] + 32;
            }
        }
        half8 rah[NQ], ral[NQ];
        half8 rbh[BD][2], rbl[BD][2];   // [set][c]: the multipole columns of BD K steps in flight
        auto aload = [&](int k0) {
#pragma unroll
            for (int j = 0; j < NQ; j++) {
                const int qi = tid + 256 * j;
                if (qi < 4 * R) {
                    const size_t o = (size_t)(qi >> 2) * Kp + k0 + (qi & 3) * 8;
                    rah[j] = *(const half8*)(Ahg + o);
                    ral[j] = *(const half8*)(Alg + o);
                }
            }
        };
        auto bload = [&](int set, int k0) {
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const half8 z = {0, 0, 0, 0, 0, 0, 0, 0};
                rbh[set][c] = bph[c] ? *(const half8*)(bph[c] + 2 * k0 + bq) : z;
                rbl[set][c] = bpl[c] ? *(const half8*)(bpl[c] + 2 * k0 + bq) : z;
            }
        };
        auto lstore = [&](int buf, int set) {
#pragma unroll
            for (int j = 0; j < NQ; j++) {
                const int qi = tid + 256 * j;
                if (qi < 4 * R) {
                    *(half8*)&Ah[buf][qi >> 2][(qi & 3) * 8] = rah[j];
                    *(half8*)&Al[buf][qi >> 2][(qi & 3) * 8] = ral[j];
                }
            }
#pragma unroll
            for (int c = 0; c < 2; c++) {
                *(half8*)&Bh[buf][bc + 64 * c][bq] = rbh[set][c];
                *(half8*)&Bl[buf][bc + 64 * c][bq] = rbl[set][c];
            }
        };
        f32x4 acc[NB0][4];
#pragma unroll
        for (int a = 0; a < NB0; a++)
#pragma unroll
            for (int c = 0; c < 4; c++) acc[a][c] = f32x4{0, 0, 0, 0};
        // the multipole columns (gathered from HBM) are loaded BD K steps ahead, the operator
        // rows (shared by the tiles of one class, mostly L2 hits) one step ahead; the K loop is
        // unrolled so the register sets are indexed statically
        aload(0);
#pragma unroll
        for (int d = 0; d < BD; d++)
            if (d < NKS) bload(d, d * BK);
        lstore(0, 0);
        __syncthreads();
#pragma unroll
        for (int kt = 0; kt < NKS; kt++) {
            const int cur = kt & 1;
            if (kt + 1 < NKS) aload((kt + 1) * BK);
            if (kt + BD < NKS) bload(kt % BD, (kt + BD) * BK);   // this set held step kt, stored already
            const int kk = 8 * g;
            half8 bh[4], bl[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                bh[c] = *(const half8*)&Bh[cur][wc * 64 + 16 * c + li][kk];
                bl[c] = *(const half8*)&Bl[cur][wc * 64 + 16 * c + li][kk];
            }
            // all of this K step's operator fragments are read up front: the MFMAs of a row block
            // wait only for its own reads, which land while the earlier blocks' MFMAs run
            half8 ah[NB0], al[NB0];
#pragma unroll
            for (int a = 0; a < NB0; a++) {
                // the second row half's last block is past NB when NB is odd: it reads the block
                // NB - 1 instead (in bounds) and its MFMAs are skipped
                const int row = min(rb0 + a, NB - 1) * 16 + li;
                ah[a] = *(const half8*)&Ah[cur][row][kk];
                al[a] = *(const half8*)&Al[cur][row][kk];
            }
#pragma unroll
            for (int a = 0; a < NB0; a++) {
                if (rb0 + a >= NB) continue;   // uniform over the wave
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], bh[c], acc[a][c], 0, 0, 0);
                    acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bl[c], acc[a][c], 0, 0, 0);
                    acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bh[c], acc[a][c], 0, 0, 0);
                }
            }
            // the other LDS buffer was last read before the previous barrier
            if (kt + 1 < NKS) lstore(cur ^ 1, (kt + 1) % BD);
            __syncthreads();
        }
        const float* rv = ops_rinv + (size_t)op * R;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int cl = wc * 64 + 16 * c + li;
            if (cl >= tile.ncol) continue;
            const int col = tile.col0 + cl, pair = col / 3, comp = col - 3 * pair;
            const float sinv = 1.0f / cscale[pair_src[pair] * 3 + comp];
            float* out = slab + (size_t)col * R;
#pragma unroll
            for (int a = 0; a < NB0; a++) {
                if (rb0 + a >= NB) break;
                const float4 r = *(const float4*)(rv + (rb0 + a) * 16 + 4 * g);
                const f32x4 v = f32x4{acc[a][c][0] * r.x * sinv, acc[a][c][1] * r.y * sinv, acc[a][c][2] * r.z * sinv,
                                      acc[a][c][3] * r.w * sinv};
                __builtin_nontemporal_store(v, (f32x4*)(out + (rb0 + a) * 16 + 4 * g));
            }
        }
    }
}

// per multipole column c (node * 3 + component): the power of 2 s_c that brings
// max_i |M[c][i] dm[i]| into [0.5, 1) (1 for a zero column), and the column's equilibrated values
// x_i = M[c][i] dm[i] s_c split into FP16 hi = fp16(x) and lo = fp16(x - hi) (zero beyond L): the
// B operand of the F16X3 M2L, split once per evaluation instead of once per tile.  Layout
// [column][K step][hi 32 | lo 32]: the hi and lo halves of one K step share a 128-byte line.
// A wave per column, lane j owning the 8 coefficients [8j, 8j + 8) (Kp <= 8 * 64): two 16-byte
// loads, a maximum over the wave, two 16-byte stores; the waves of a capped grid stride over the
// columns and keep their lanes' dm in registers.
__global__ __launch_bounds__(256) void k_split_columns(const float* __restrict__ M, int ncols, int L, int Lp, int Kp,
                                                       const float* __restrict__ dm, float* __restrict__ scale,
                                                       _Float16* __restrict__ hl,
                                                       const int* __restrict__ skip_first_child,
                                                       const int* __restrict__ skip_npts) {
    const int lane = threadIdx.x & 63, i0 = 8 * lane;
    const bool act = i0 < Kp;   // Kp and Lp are multiples of 8, Kp <= Lp
    const int off = (i0 >> 5) * 64 + (i0 & 31);   // K step i0 / 32, chunk within the step
    float d[8];
#pragma unroll
    for (int e = 0; e < 8; e++) d[e] = act && i0 + e < L ? dm[i0 + e] : 0.0f;
    for (int c = blockIdx.x * 4 + (threadIdx.x >> 6); c < ncols; c += gridDim.x * 4) {
        if (skip_first_child) {   // leaves with points: split by the P2M already
            const int node = c / 3;
            if (skip_first_child[node] < 0 && skip_npts[node] > 0) continue;
        }
        float x[8];
        float mx = 0.0f;
        if (act) {
            const float4 u = *(const float4*)(M + (size_t)c * Lp + i0);
            const float4 v = *(const float4*)(M + (size_t)c * Lp + i0 + 4);
            x[0] = u.x * d[0]; x[1] = u.y * d[1]; x[2] = u.z * d[2]; x[3] = u.w * d[3];
            x[4] = v.x * d[4]; x[5] = v.y * d[5]; x[6] = v.z * d[6]; x[7] = v.w * d[7];
#pragma unroll
            for (int e = 0; e < 8; e++) mx = fmaxf(mx, fabsf(x[e]));
        }
        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        const float sc = mx > 0.0f ? exp2f((float)(-(ilogbf(mx) + 1))) : 1.0f;
        if (lane == 0) scale[c] = sc;
        if (act) {
            half8 h, l;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const float y = x[e] * sc;
                h[e] = (_Float16)y;
                l[e] = (_Float16)(y - (float)h[e]);
            }
            _Float16* dst = hl + (size_t)c * 2 * Kp + off;
            *(half8*)dst = h;
            *(half8*)(dst + 32) = l;
        }
    }
}

void split_columns(const float* M, int ncols, int L, int Lp, int Kp, const float* dm, float* scale, _Float16* hl,
                   hipStream_t st, const int* skip_first_child, const int* skip_npts) {
    if (ncols <= 0) return;
    if (Kp > 8 * 64 || Kp % 32 || Lp % 8 || Kp > Lp) throw Error(4, "FP16 split: column length out of range");
    if (!skip_first_child != !skip_npts) throw Error(1, "FP16 split: node flags come in pairs");
    const int grid = std::min((ncols + 3) / 4, 8 * num_cus());
    k_split_columns<<<grid, 256, 0, st>>>(M, ncols, L, Lp, Kp, dm, scale, hl, skip_first_child, skip_npts);
    FMM_HIP(hipGetLastError());
}

int m2l_f16_rows(int L) { return 16 * ((L + 15) / 16); }
bool m2l_f16x3_persistent(int L) {
    const int nb = m2l_f16_rows(L) / 16;
    return nb >= 14 && nb <= 23;   // the 128-column tiles of launch_f16x3_128 below
}
int m2l_f16_kpad(int L) { return 32 * ((L + 31) / 32); }

// 128-column tiles: the double-buffered 2 x 2 kernel where the single-buffered one's
// accumulators leave one workgroup per CU anyway (R >= 224, p >= 13) and its two LDS buffers fit
// (R <= 368, p <= 18); the single-buffered kernel (two or more workgroups per CU below) otherwise
template <int NB, int NKS>
static void launch_f16x3_128(int ntiles, const _Float16* ops_hi, const _Float16* ops_lo, const float* ops_rinv,
                             const GemmTile* d_tiles, const int* pair_src, const _Float16* src_hl,
                             const float* cscale, float* slab, int op_base, int* queues, hipStream_t st) {
    if constexpr (NB >= 14 && NB <= 23) {
        // one workgroup per CU (the two LDS buffers), 8 per XCD round
        const int grid = std::min(8 * ((ntiles + 7) / 8), 8 * ((num_cus() + 7) / 8));
        FMM_HIP(hipMemsetAsync(queues, 0, 8 * sizeof(int), st));
        k_gemm_m2l_f16x3_q<NB, NKS, 2><<<grid, 256, 0, st>>>(ops_hi, ops_lo, ops_rinv, d_tiles, pair_src, src_hl,
                                                             cscale, slab, op_base, ntiles, queues);
    }
    else
        k_gemm_m2l_f16x3<NB, NKS, 2><<<ntiles, 256, 0, st>>>(ops_hi, ops_lo, ops_rinv, d_tiles, pair_src, src_hl,
                                                             cscale, slab, op_base);
}

void m2l_gemm_f16x3(const _Float16* ops_hi, const _Float16* ops_lo, const float* ops_rinv, int op_base,
                    const GemmTile* d_tiles, int ntiles, int tile_cols, const int* pair_src, const _Float16* src_hl,
                    const float* cscale, float* slab, int L, int* queues, hipStream_t st) {
    if (ntiles == 0) return;
    const int nb = m2l_f16_rows(L) / 16, nks = m2l_f16_kpad(L) / 32;
    if (tile_cols != 64 && tile_cols != 128) throw Error(4, "FP16 M2L: tiles of 64 or 128 columns");
#define FMM_F16(NBV, NKV)                                                                                            \
    if (nb == NBV && nks == NKV) {                                                                                   \
        if (tile_cols == 128)                                                                                        \
            launch_f16x3_128<NBV, NKV>(ntiles, ops_hi, ops_lo, ops_rinv, d_tiles, pair_src, src_hl, cscale, slab,    \
                                       op_base, queues, st);                                                         \
        else                                                                                                         \
            k_gemm_m2l_f16x3<NBV, NKV, 1><<<ntiles, 256, 0, st>>>(ops_hi, ops_lo, ops_rinv, d_tiles, pair_src,       \
                                                                  src_hl, cscale, slab, op_base);                    \
        FMM_HIP(hipGetLastError());                                                                                  \
        return;                                                                                                      \
    }
    FMM_F16(1, 1) FMM_F16(2, 1) FMM_F16(3, 2) FMM_F16(4, 2) FMM_F16(4, 3) FMM_F16(5, 3) FMM_F16(6, 3) FMM_F16(7, 4)
    FMM_F16(8, 4) FMM_F16(9, 5) FMM_F16(10, 5) FMM_F16(11, 6) FMM_F16(12, 6) FMM_F16(13, 7) FMM_F16(14, 7)
    FMM_F16(15, 8) FMM_F16(16, 8) FMM_F16(17, 9) FMM_F16(18, 9) FMM_F16(19, 10) FMM_F16(20, 10) FMM_F16(21, 11)
    FMM_F16(22, 11) FMM_F16(23, 12) FMM_F16(24, 12) FMM_F16(25, 13) FMM_F16(26, 13) FMM_F16(27, 14) FMM_F16(28, 14)
#undef FMM_F16
    throw Error(4, "FP16 M2L: no kernel instance for this order");
}

template <int NB>
static void launch_m2l_nb(Precision prec, const void* ops, const GemmTile* d_tiles, int ntiles, const int* pair_src,
                          const void* src_exp, void* slab, int Lp, hipStream_t st) {
    if (prec == Precision::F64)
        k_gemm_m2l<double, NB><<<ntiles, 256, 0, st>>>((const double*)ops, d_tiles, pair_src, (const double*)src_exp,
                                                       (double*)slab, Lp);
    else
        k_gemm_m2l<float, NB><<<ntiles, 256, 0, st>>>((const float*)ops, d_tiles, pair_src, (const float*)src_exp,
                                                      (float*)slab, Lp);
    FMM_HIP(hipGetLastError());
}

int m2l_slab_rows(Precision prec, int L, int Lp) {
    if (prec == Precision::BF16X3) return Lp;
    if (prec == Precision::F16X3) return m2l_f16_rows(L);
    int nb = (L + 15) / 16;
    switch (nb) {
        case 1: case 2: case 3: case 4: case 6: case 7: case 8: case 9: case 11: case 13: case 15: case 16: case 19:
        case 21: case 23: case 25: case 28:
            return 16 * nb;
        default:
            return Lp;
    }
}

void m2l_gemm(Precision prec, const void* ops, const void* ops_lo, const GemmTile* d_tiles, int ntiles,
              const int* pair_src, const void* src_exp, void* slab, int L, int Lp, hipStream_t st) {
    if (ntiles == 0) return;
    const int R = m2l_slab_rows(prec, L, Lp);
    if (R == Lp) {   // the whole-panel kernel (BF16x3, or an L without a padded-to-16 instance)
        class_gemm(prec, ops, ops_lo, d_tiles, ntiles, pair_src, nullptr, src_exp, slab, Lp, EPI_STORE, st);
        return;
    }
#define FMM_NB(N) \
    case N: launch_m2l_nb<N>(prec, ops, d_tiles, ntiles, pair_src, src_exp, slab, Lp, st); break;
    switch (R / 16) {
        FMM_NB(1) FMM_NB(2) FMM_NB(3) FMM_NB(4) FMM_NB(6) FMM_NB(7) FMM_NB(8) FMM_NB(9) FMM_NB(11) FMM_NB(13)
        FMM_NB(15) FMM_NB(16) FMM_NB(19) FMM_NB(21) FMM_NB(23) FMM_NB(25) FMM_NB(28)
        default: throw Error(4, "no padded-row M2L kernel for this order");
    }
#undef FMM_NB
}

template <int RB, bool TR, bool MU>
static void launch_rb(Precision prec, const void* ops, const void* ops_lo, const GemmTile* d_tiles, int ntiles,
                      const int* pair_src, const int* pair_dst, const void* src_exp, void* dst_exp, int epi,
                      int nseg, int tri, hipStream_t st) {
    if (prec == Precision::F64)
        k_gemm<double, RB, TR, MU><<<ntiles, 256, 0, st>>>((const double*)ops, d_tiles, pair_src, pair_dst,
                                                          (const double*)src_exp, (double*)dst_exp, epi, nseg, tri);
    else if (prec == Precision::F32)
        k_gemm<float, RB, TR, MU><<<ntiles, 256, 0, st>>>((const float*)ops, d_tiles, pair_src, pair_dst,
                                                         (const float*)src_exp, (float*)dst_exp, epi, nseg, tri);
    else
        k_gemm_bf16x3<RB, TR, MU><<<ntiles, 256, 0, st>>>((const __bf16*)ops, (const __bf16*)ops_lo, d_tiles, pair_src,
                                                         pair_dst, (const float*)src_exp, (float*)dst_exp, epi, nseg);
    FMM_HIP(hipGetLastError());
}

template <bool TR>
static void launch(int RB, Precision prec, const void* ops, const void* ops_lo, const GemmTile* d_tiles, int ntiles,
                   const int* pair_src, const int* pair_dst, const void* src_exp, void* dst_exp, int epi, int nseg,
                   int tri, hipStream_t st) {
#define FMM_RB(R)                                                                                             \
    case R:                                                                                                   \
        if (nseg > 1)                                                                                         \
            launch_rb<R, TR, true>(prec, ops, ops_lo, d_tiles, ntiles, pair_src, pair_dst, src_exp, dst_exp, epi, \
                                   nseg, tri, st);                                                            \
        else                                                                                                  \
            launch_rb<R, TR, false>(prec, ops, ops_lo, d_tiles, ntiles, pair_src, pair_dst, src_exp, dst_exp, epi, \
                                    1, tri, st);                                                              \
        break;
    switch (RB) {
        FMM_RB(1) FMM_RB(2) FMM_RB(3) FMM_RB(4) FMM_RB(5) FMM_RB(6) FMM_RB(7)
        default: throw Error(4, "padded expansion length above 448 (order > 20)");
    }
#undef FMM_RB
}

void class_gemm(Precision prec, const void* ops, const void* ops_lo, const GemmTile* d_tiles, int ntiles,
                const int* pair_src, const int* pair_dst, const void* src_exp, void* dst_exp, int Lp, int epilogue,
                hipStream_t st, bool transpose, int nseg, int tri) {
    if (ntiles == 0) return;
    if (tri < TRI_NONE || tri > TRI_UPPER) throw Error(1, "unknown operator structure");
    if (Lp % 64 != 0) throw Error(1, "padded expansion length must be a multiple of 64");
    int RB = Lp / 64;
    if (epilogue < EPI_ADD || epilogue > EPI_STORE) throw Error(1, "unknown GEMM epilogue");
    if (nseg < 1) throw Error(1, "a GEMM tile needs at least one segment");
    if (transpose)
        launch<true>(RB, prec, ops, ops_lo, d_tiles, ntiles, pair_src, pair_dst, src_exp, dst_exp, epilogue, nseg, tri,
                     st);
    else
        launch<false>(RB, prec, ops, ops_lo, d_tiles, ntiles, pair_src, pair_dst, src_exp, dst_exp, epilogue, nseg, tri,
                      st);
}

// host-side tiling of class column ranges: class c owns pairs [begin[c], begin[c+1])
void plan_tiles(const std::vector<int>& class_begin, const std::vector<int>& class_op, std::vector<GemmTile>& tiles,
                int width) {
    tiles.clear();
    int nc = (int)class_begin.size() - 1;
    for (int c = 0; c < nc; c++) {
        int c0 = 3 * class_begin[c], c1 = 3 * class_begin[c + 1];
        for (int j = c0; j < c1; j += width) tiles.push_back(GemmTile{class_op[c], j, std::min(width, c1 - j)});
    }
}

// ---------------------------------------------------------------- per-node reduction of pair slabs
__global__ void k_iota(int* v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}
__global__ void k_zero_ranges(int* b, int* e, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = e[i] = 0;
}
__global__ void k_key_ranges(const uint32_t* __restrict__ sk, int n, int* __restrict__ b, int* __restrict__ e) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = sk[i];
    if (i == 0 || sk[i - 1] != k) b[k] = i;
    if (i == n - 1 || sk[i + 1] != k) e[k] = i + 1;
}

__global__ void k_segment_counts(const int* __restrict__ b, const int* __restrict__ e, int n, uint32_t* __restrict__ cnt) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) cnt[v] = (uint32_t)(e[v] - b[v] + PAIR_SEGMENT - 1) / PAIR_SEGMENT;
    else if (v == n) cnt[v] = 0;
}
__global__ void k_segment_nodes(const uint32_t* __restrict__ off, int n, int* __restrict__ seg_node) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    for (uint32_t s = off[v]; s < off[v + 1]; s++) seg_node[s] = v;
}

static int key_bits(int nnodes) {
    int b = 1;
    while (b < 32 && (1u << b) < (unsigned)nnodes) b++;
    return b;
}

void build_pair_index(const int* key, int npairs, int nnodes, PairIndex& ix, hipStream_t st) {
    ix.npairs = npairs;
    ix.nnodes = nnodes;
    int* b = ix.begin.get<int>(std::max(nnodes, 1));
    int* e = ix.end.get<int>(std::max(nnodes, 1));
    if (nnodes) k_zero_ranges<<<blocks(nnodes), 256, 0, st>>>(b, e, nnodes);
    if (!npairs) return;
    int* io = ix.iota.get<int>(npairs);
    k_iota<<<blocks(npairs), 256, 0, st>>>(io, npairs);
    uint32_t* ks = ix.keys_sorted.get<uint32_t>(npairs);
    int* ord = ix.order.get<int>(npairs);
    // stable radix sort: pairs of one node stay in ascending pair order
    sort_pairs_u32((const uint32_t*)key, ks, io, ord, npairs, key_bits(nnodes), ix.tmp, st);
    k_key_ranges<<<blocks(npairs), 256, 0, st>>>(ks, npairs, b, e);
    // segments of at most PAIR_SEGMENT pairs per node, numbered node by node
    uint32_t* cnt = ix.seg_count.get<uint32_t>(nnodes + 1);
    uint32_t* off = ix.seg_off.get<uint32_t>(nnodes + 1);
    k_segment_counts<<<blocks(nnodes + 1), 256, 0, st>>>(b, e, nnodes, cnt);
    exclusive_scan_u32(cnt, off, nnodes + 1, ix.tmp, st);
    ix.max_segments = npairs / PAIR_SEGMENT + std::min(npairs, nnodes);
    k_segment_nodes<<<blocks(nnodes), 256, 0, st>>>(off, nnodes, ix.seg_node.get<int>(ix.max_segments));
    FMM_HIP(hipGetLastError());
}

// Two passes, so that nodes with hundreds of pairs do not serialise the launch: (1) one wave
// per segment (at most PAIR_SEGMENT consecutive pairs of one node) sums its pairs' slab blocks
// in ascending pair order into a partial block; (2) one wave per node sums its segments'
// partials, in segment order, and writes the node's expansion.  The order of every sum is fixed
// by the pair numbering alone (deterministic).  A pair's slab block is nv = 3 vr 16-byte vectors (vr per
// component: the first R of each component's Lp coefficients); lane l owns the vectors
// l + 64 r, r < NR = ceil(nv / 64); pass (1) keeps the loads of U pairs in flight.
__device__ __forceinline__ float4 zero_vec(float4) { return make_float4(0.f, 0.f, 0.f, 0.f); }
__device__ __forceinline__ double2 zero_vec(double2) { return make_double2(0.0, 0.0); }
// streaming loads of the slab (read once)
__device__ __forceinline__ float4 load_stream(const float4* p) {
    f32x4 v = __builtin_nontemporal_load((const f32x4*)p);
    return make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ double2 load_stream(const double2* p) {
    typedef double f64x2 __attribute__((ext_vector_type(2)));
    f64x2 v = __builtin_nontemporal_load((const f64x2*)p);
    return make_double2(v[0], v[1]);
}

template <class V, int NR>
__global__ __launch_bounds__(256) void k_reduce_segments(const uint32_t* __restrict__ seg_off,
                                                         const int* __restrict__ seg_node, const int* __restrict__ begin,
                                                         const int* __restrict__ end, const int* __restrict__ order,
                                                         const V* __restrict__ slab, int nnodes, int nv,
                                                         V* __restrict__ partial) {
    constexpr int U = NR <= 4 ? 4 : NR <= 8 ? 2 : 1;   // 8 to 16 vectors in flight per lane
    const int sg = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (sg >= (int)seg_off[nnodes]) return;
    const int v = seg_node[sg];
    const int k0 = begin[v] + (sg - (int)seg_off[v]) * PAIR_SEGMENT, k1 = min(k0 + PAIR_SEGMENT, end[v]);
    V acc[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) acc[r] = zero_vec(V());
    int k = k0;
    for (; k + U <= k1; k += U) {
        V x[U][NR];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const V* src = slab + (size_t)order[k + u] * nv + lane;
#pragma unroll
            for (int r = 0; r < NR; r++)
                if (lane + 64 * r < nv) x[u][r] = load_stream(src + 64 * r);
        }
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int r = 0; r < NR; r++)
                if (lane + 64 * r < nv) acc[r] = acc[r] + x[u][r];
    }
    for (; k < k1; k++) {
        const V* src = slab + (size_t)order[k] * nv + lane;
#pragma unroll
        for (int r = 0; r < NR; r++)
            if (lane + 64 * r < nv) acc[r] = acc[r] + load_stream(src + 64 * r);
    }
    V* o = partial + (size_t)sg * nv + lane;
#pragma unroll
    for (int r = 0; r < NR; r++)
        if (lane + 64 * r < nv) o[64 * r] = acc[r];
}

template <class V, int NR>
__global__ __launch_bounds__(256) void k_reduce_nodes(const uint32_t* __restrict__ seg_off, const V* __restrict__ partial,
                                                      int nnodes, int vr, int vl, V* __restrict__ out) {
    const int v = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (v >= nnodes) return;
    const int s0 = (int)seg_off[v], s1 = (int)seg_off[v + 1];
    const int nv = 3 * vr;
    V* o = out + (size_t)v * 3 * vl;
    // the node's expansion is written whole (nothing clears it beforehand): the padding vectors
    // vr .. vl - 1 of each component, and all of a node without far pairs, are zero
    for (int i = lane; i < 3 * vl; i += 64) {
        const int c = i / vl;
        if (s0 == s1 || i - c * vl >= vr) o[i] = zero_vec(V());
    }
    if (s0 == s1) return;
    // the loads of four segments are issued together; the adds keep the segment order
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int i = lane + 64 * r;
        if (i >= nv) continue;
        const int c = i / vr, q = c * vl + (i - c * vr);
        V acc = zero_vec(V());
        const V* pp = partial + (size_t)s0 * nv + i;
        int sg = s0;
        for (; sg + 4 <= s1; sg += 4, pp += 4 * (size_t)nv) {
            const V x0 = pp[0], x1 = pp[nv], x2 = pp[2 * (size_t)nv], x3 = pp[3 * (size_t)nv];
            acc = acc + x0;
            acc = acc + x1;
            acc = acc + x2;
            acc = acc + x3;
        }
        for (; sg < s1; sg++, pp += nv) acc = acc + pp[0];
        o[q] = acc;
    }
}

template <class V>
static void launch_reduce(PairIndex& ix, const void* slab, int vr, int vl, void* out, hipStream_t st) {
    const int nv = 3 * vr, nr = (nv + 63) / 64;
    const uint32_t* off = ix.seg_off.as<uint32_t>();
    const int* sn = ix.seg_node.as<int>();
    V* part = (V*)ix.partial.reserve((size_t)std::max(ix.max_segments, 1) * nv * sizeof(V));
    const int g1 = (std::max(ix.max_segments, 1) + 3) / 4, g2 = (ix.nnodes + 3) / 4;
    const int *b = ix.begin.as<int>(), *e = ix.end.as<int>(), *o = ix.order.as<int>();
#define FMM_NR(N)                                                                                            \
    case N:                                                                                                 \
        k_reduce_segments<V, N><<<g1, 256, 0, st>>>(off, sn, b, e, o, (const V*)slab, ix.nnodes, nv, part); \
        k_reduce_nodes<V, N><<<g2, 256, 0, st>>>(off, part, ix.nnodes, vr, vl, (V*)out);                   \
        break;
    switch (nr) {
        FMM_NR(1) FMM_NR(2) FMM_NR(3) FMM_NR(4) FMM_NR(5) FMM_NR(6) FMM_NR(7) FMM_NR(8) FMM_NR(9) FMM_NR(10)
        FMM_NR(11) FMM_NR(12) FMM_NR(13) FMM_NR(14) FMM_NR(15) FMM_NR(16)
        default: throw Error(4, "slab rows too long for the pair reduction");
    }
#undef FMM_NR
    FMM_HIP(hipGetLastError());
}

void reduce_pairs(bool f64, PairIndex& ix, const void* slab, int R, int Lp, void* out, hipStream_t st) {
    if (!ix.npairs || !ix.nnodes) return;
    if (f64) launch_reduce<double2>(ix, slab, R / 2, Lp / 2, out, st);
    else launch_reduce<float4>(ix, slab, R / 4, Lp / 4, out, st);
}

}  // namespace fmm

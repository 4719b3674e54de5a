// Class-batched translation GEMMs on MFMA (gfx950).
//
// A "class" is one translation operator T (Lp x Lp, row = output coefficient) shared
// by a list of (source node, destination node) pairs: an M2L offset class, or one M2M /
// L2L orientation of a level.  Column j of a class is component j % 3 of pair j / 3:
// B[:, j] = src_exp[src][comp][:], and the product T B[:, j] is added to
// dst_exp[dst][comp][:].  One launch covers every column tile of every class; the
// column tiles never straddle classes.  Expansions are stored node-major as
// [node][3][Lp] so a column is one contiguous, 16-byte-aligned vector.
//
// Precisions (template Mode):
//   F64    : v_mfma_f64_16x16x4f64, FP64 operators and expansions;
//   F32    : v_mfma_f32_16x16x4f32, FP32 operators/expansions, FP32 accumulate;
//   BF16X3 : v_mfma_f32_16x16x32_bf16 on split operands A = Ah + Al (pre-split tables),
//            B = Bh + Bl (split on load), C += Ah Bh + Ah Bl + Al Bh, FP32 accumulate.
// Tile: 64 rows x 64 columns per 256-thread workgroup; each wave owns a 32 x 32 block.
#include "fmm_gemm.hpp"

namespace fmm {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int BM = 64, BN = 64;

__device__ __forceinline__ void add_out(float* p, float v, int atomic) {
    if (atomic) atomicAdd(p, v);
    else *p += v;
}
__device__ __forceinline__ void add_out(double* p, double v, int atomic) {
    if (atomic) atomicAdd(p, v);
    else *p += v;
}

// ---------------------------------------------------------------- FP64 / FP32
template <class T, bool TRANS>
__global__ __launch_bounds__(256) void k_class_gemm(const T* __restrict__ ops, const GemmTile* __restrict__ tiles,
                                                    const int* __restrict__ pair_src, const int* __restrict__ pair_dst,
                                                    const T* __restrict__ src_exp, T* __restrict__ dst_exp, int Lp,
                                                    int atomic) {
    constexpr int BK = 16;
    __shared__ T As[BK][BM + 1];
    __shared__ T Bs[BK][BN + 1];
    const GemmTile tile = tiles[blockIdx.x];
    const int row0 = blockIdx.y * BM;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const T* A = ops + (size_t)tile.op * Lp * Lp;
    // loader coordinates: 256 threads x 4 elements for each of A (64x16) and B (16x64)
    const int lr = tid >> 2, lk = (tid & 3) * 4;
    const int arow = row0 + lr;
    const int bcol = tile.col0 + lr;
    const bool bvalid = lr < tile.ncol;
    const T* bptr = nullptr;
    if (bvalid) {
        int pair = bcol / 3, comp = bcol - 3 * pair;
        bptr = src_exp + ((size_t)pair_src[pair] * 3 + comp) * Lp;
    }
    typedef typename std::conditional<sizeof(T) == 8, f64x4, f32x4>::type acc_t;
    acc_t acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = acc_t{0, 0, 0, 0};
    for (int k0 = 0; k0 < Lp; k0 += BK) {
        for (int q = 0; q < 4; q++) {
            size_t ai = TRANS ? (size_t)(k0 + lk + q) * Lp + arow : (size_t)arow * Lp + k0 + lk + q;
            As[lk + q][lr] = (arow < Lp) ? A[ai] : T(0);
            Bs[lk + q][lr] = bvalid ? bptr[k0 + lk + q] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            T af[2], bf[2];
            for (int a = 0; a < 2; a++) af[a] = As[ks + (lane >> 4)][wm + a * 16 + (lane & 15)];
            for (int b = 0; b < 2; b++) bf[b] = Bs[ks + (lane >> 4)][wn + b * 16 + (lane & 15)];
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) {
                    if constexpr (sizeof(T) == 8)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
                    else
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], bf[b], acc[a][b], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    // epilogue: scatter-add into the destination expansions
    for (int b = 0; b < 2; b++) {
        int cl = wn + b * 16 + (lane & 15);
        if (cl >= tile.ncol) continue;
        int col = tile.col0 + cl;
        int pair = col / 3, comp = col - 3 * pair;
        T* dst = dst_exp + ((size_t)pair_dst[pair] * 3 + comp) * Lp;
        for (int a = 0; a < 2; a++)
            for (int r = 0; r < 4; r++) {
                // f64 MFMA: reg r holds row 4r + lane/16; f32 MFMA: row 4(lane/16) + r
                int rr = (sizeof(T) == 8) ? (4 * r + (lane >> 4)) : (4 * (lane >> 4) + r);
                int row = row0 + wm + a * 16 + rr;
                if (row < Lp) add_out(dst + row, (T)acc[a][b][r], atomic);
            }
    }
}

// ---------------------------------------------------------------- BF16x3
__device__ __forceinline__ void split_bf16(float v, __bf16& h, __bf16& l) {
    h = (__bf16)v;
    l = (__bf16)(v - (float)h);
}

template <bool TRANS>
__global__ __launch_bounds__(256) void k_class_gemm_bf16x3(const __bf16* __restrict__ ops_hi,
                                                           const __bf16* __restrict__ ops_lo,
                                                           const GemmTile* __restrict__ tiles,
                                                           const int* __restrict__ pair_src,
                                                           const int* __restrict__ pair_dst,
                                                           const float* __restrict__ src_exp,
                                                           float* __restrict__ dst_exp, int Lp, int atomic) {
    constexpr int BK = 32;
    // k-contiguous rows so each lane reads its 8 consecutive k values as one vector
    __shared__ __attribute__((aligned(16))) __bf16 Ah[BM][BK + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Al[BM][BK + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Bh[BN][BK + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Bl[BN][BK + 8];
    const GemmTile tile = tiles[blockIdx.x];
    const int row0 = blockIdx.y * BM;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const size_t opoff = (size_t)tile.op * Lp * Lp;
    const int lr = tid >> 2, lk = (tid & 3) * 8;   // 256 threads x 8 elements = 64 x 32
    const int arow = row0 + lr;
    const int bcol = tile.col0 + lr;
    const bool bvalid = lr < tile.ncol;
    const float* bptr = nullptr;
    if (bvalid) {
        int pair = bcol / 3, comp = bcol - 3 * pair;
        bptr = src_exp + ((size_t)pair_src[pair] * 3 + comp) * Lp;
    }
    f32x4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0, 0, 0, 0};
    for (int k0 = 0; k0 < Lp; k0 += BK) {
        for (int q = 0; q < 8; q++) {
            int k = k0 + lk + q;
            bool ain = arow < Lp && k < Lp;
            size_t ai = opoff + (TRANS ? (size_t)k * Lp + arow : (size_t)arow * Lp + k);
            Ah[lr][lk + q] = ain ? ops_hi[ai] : (__bf16)0.0f;
            Al[lr][lk + q] = ain ? ops_lo[ai] : (__bf16)0.0f;
            float v = (bvalid && k < Lp) ? bptr[k] : 0.0f;
            __bf16 h, l;
            split_bf16(v, h, l);
            Bh[lr][lk + q] = h;
            Bl[lr][lk + q] = l;
        }
        __syncthreads();
        bf16x8 ah[2], al[2], bh[2], bl[2];
        const int kk = 8 * (lane >> 4);
        for (int a = 0; a < 2; a++) {
            int r = wm + a * 16 + (lane & 15);
            ah[a] = *(const bf16x8*)&Ah[r][kk];
            al[a] = *(const bf16x8*)&Al[r][kk];
        }
        for (int b = 0; b < 2; b++) {
            int c = wn + b * 16 + (lane & 15);
            bh[b] = *(const bf16x8*)&Bh[c][kk];
            bl[b] = *(const bf16x8*)&Bl[c][kk];
        }
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
            }
        __syncthreads();
    }
    for (int b = 0; b < 2; b++) {
        int cl = wn + b * 16 + (lane & 15);
        if (cl >= tile.ncol) continue;
        int col = tile.col0 + cl;
        int pair = col / 3, comp = col - 3 * pair;
        float* dst = dst_exp + ((size_t)pair_dst[pair] * 3 + comp) * Lp;
        for (int a = 0; a < 2; a++)
            for (int r = 0; r < 4; r++) {
                int row = row0 + wm + a * 16 + 4 * (lane >> 4) + r;
                if (row < Lp) add_out(dst + row, acc[a][b][r], atomic);
            }
    }
}

template <bool TR>
static void launch(Precision prec, const void* ops, const void* ops_lo, const GemmTile* d_tiles, int ntiles,
                   const int* pair_src, const int* pair_dst, const void* src_exp, void* dst_exp, int Lp, bool atomic,
                   hipStream_t st) {
    dim3 grid(ntiles, (Lp + BM - 1) / BM);
    if (prec == Precision::F64)
        k_class_gemm<double, TR><<<grid, 256, 0, st>>>((const double*)ops, d_tiles, pair_src, pair_dst,
                                                       (const double*)src_exp, (double*)dst_exp, Lp, atomic ? 1 : 0);
    else if (prec == Precision::F32)
        k_class_gemm<float, TR><<<grid, 256, 0, st>>>((const float*)ops, d_tiles, pair_src, pair_dst,
                                                      (const float*)src_exp, (float*)dst_exp, Lp, atomic ? 1 : 0);
    else
        k_class_gemm_bf16x3<TR><<<grid, 256, 0, st>>>((const __bf16*)ops, (const __bf16*)ops_lo, d_tiles, pair_src,
                                                     pair_dst, (const float*)src_exp, (float*)dst_exp, Lp,
                                                     atomic ? 1 : 0);
    FMM_HIP(hipGetLastError());
}

void class_gemm(Precision prec, const void* ops, const void* ops_lo, const GemmTile* d_tiles, int ntiles,
                const int* pair_src, const int* pair_dst, const void* src_exp, void* dst_exp, int Lp, bool atomic,
                hipStream_t st, bool transpose) {
    if (ntiles == 0) return;
    if (Lp % 32 != 0) throw Error(1, "padded expansion length must be a multiple of 32");
    if (transpose) launch<true>(prec, ops, ops_lo, d_tiles, ntiles, pair_src, pair_dst, src_exp, dst_exp, Lp, atomic, st);
    else launch<false>(prec, ops, ops_lo, d_tiles, ntiles, pair_src, pair_dst, src_exp, dst_exp, Lp, atomic, st);
}

// host-side tiling of class column ranges: class c owns pairs [begin[c], begin[c+1])
void plan_tiles(const std::vector<int>& class_begin, const std::vector<int>& class_op, std::vector<GemmTile>& tiles) {
    tiles.clear();
    int nc = (int)class_begin.size() - 1;
    for (int c = 0; c < nc; c++) {
        int c0 = 3 * class_begin[c], c1 = 3 * class_begin[c + 1];
        for (int j = c0; j < c1; j += BN) tiles.push_back(GemmTile{class_op[c], j, std::min(BN, c1 - j)});
    }
}

}  // namespace fmm

#pragma once
#include "fmm_common.hpp"

namespace fmm {

enum class Precision { F64 = 0, F32 = 1, BF16X3 = 2, F16X3 = 3 };

struct GemmTile {
    int op;      // operator index
    int col0;    // first global column (3 * pair + component)
    int ncol;    // valid columns in this tile (<= 64)
};

// How a class GEMM delivers column j = 3 * pair + component of T B:
enum Epilogue {
    EPI_ADD = 0,      // dst_exp[pair_dst[pair]][component] += (plain; each destination once per launch)
    EPI_ATOMIC = 1,   // the same with hardware FP atomics (destinations repeat)
    EPI_STORE = 2     // written to dst_exp + j * Lp (a per-pair slab; pair_dst unused)
};
// Degree structure of the applied operator (row = output coefficient of degree l, column = input
// coefficient of degree l'): TRI_LOWER has zeros wherever l' > l (M2M, and the transposed L2L),
// TRI_UPPER wherever l' < l (L2L, and the transposed M2M).  The FP64/FP32 GEMM skips the 16 x 16
// blocks that are zero by this structure.
enum { TRI_NONE = 0, TRI_LOWER = 1, TRI_UPPER = 2 };

// ops: operator tables (double / float / bf16 hi), ops_lo: bf16 lo parts (BF16X3 only).
// transpose: apply T^T (adjoint passes).
// src_exp / dst_exp: [node][3][Lp] in double (F64) or float (F32, BF16X3).
// nseg > 1: every column sums nseg products, segment g applying operator tile.op + g to
// node pair_src + g (the 8 children of a parent, first child in pair_src).
void class_gemm(Precision prec, const void* ops, const void* ops_lo, const GemmTile* d_tiles, int ntiles,
                const int* pair_src, const int* pair_dst, const void* src_exp, void* dst_exp, int Lp, int epilogue,
                hipStream_t st, bool transpose = false, int nseg = 1, int tri = TRI_NONE);

// Pairs grouped by the node that receives them: order[begin[v] .. end[v]) are the pair indices
// (ascending) whose key is v; nodes without pairs have begin = end = 0.  Each node's range is
// cut into segments of at most PAIR_SEGMENT pairs, numbered node by node: node v owns segments
// [seg_off[v], seg_off[v + 1]), seg_node maps a segment to its node (max_segments bounds their
// number on the host without a read-back); partial holds one reduced block per segment.
constexpr int PAIR_SEGMENT = 32;
struct PairIndex {
    int npairs = 0, nnodes = 0, max_segments = 0;
    DevBuf keys, keys_sorted, iota, order, begin, end, tmp, seg_count, seg_off, seg_node, partial;
};
void build_pair_index(const int* key, int npairs, int nnodes, PairIndex& ix, hipStream_t st);

// out[v][c][0 .. R) = sum over the pairs of v, in ascending pair order, of slab[pair][c][0 .. R),
// out[v][c][R .. Lp) = 0, for every node v < ix.nnodes (written, not added: nothing needs to clear
// out first); slab = a per-pair store with row length R <= Lp, multiple of 4; double if f64 else
// float
void reduce_pairs(bool f64, PairIndex& ix, const void* slab, int R, int Lp, void* out, hipStream_t st);

// Row length of the M2L slab: L padded to 16 (MFMA block) where a kernel instance exists for it
// (FP32/FP64, F16X3), else Lp (BF16X3, whose K step is 32, and orders without an instance).
int m2l_slab_rows(Precision prec, int L, int Lp);
// M2L of all classes into the slab [3 * pairs][m2l_slab_rows] (column 3 pair + component).
void m2l_gemm(Precision prec, const void* ops, const void* ops_lo, const GemmTile* d_tiles, int ntiles,
              const int* pair_src, const void* src_exp, void* slab, int L, int Lp, hipStream_t st);

// F16X3 M2L (FP16 split products on MFMA, equilibrated by powers of 2): operator rows
// m2l_f16_rows(L), K padded to Kp = m2l_f16_kpad(L); the multipoles enter split by split_columns
// (src_hl [node * 3 + comp][Kp / 32][hi 32 | lo 32], equilibrated by dm[>= Kp], the per-coefficient
// scales the tables were built with, and by cscale[node * 3 + comp]); the slab rows are
// m2l_f16_rows(L) long.  Tile operator k is table entry k - op_base; tiles hold at most tile_cols
// (64 or 128) columns.  queues: 8 device ints, the per-XCD tile counters of the persistent kernel.
int m2l_f16_rows(int L);
// whether the F16X3 M2L of order (L = (p + 1)^2) runs as the persistent one-workgroup-per-CU kernel
// (128-column tiles) that leaves room for a near-field workgroup per CU
bool m2l_f16x3_persistent(int L);
int m2l_f16_kpad(int L);
// skip_first_child / skip_npts (per node, both or neither): columns of leaves with points are
// left as they are (the P2M writes them, see launch_p2m)
void split_columns(const float* M, int ncols, int L, int Lp, int Kp, const float* dm, float* scale, _Float16* hl,
                   hipStream_t st, const int* skip_first_child = nullptr, const int* skip_npts = nullptr);
void m2l_gemm_f16x3(const _Float16* ops_hi, const _Float16* ops_lo, const float* ops_rinv, int op_base,
                    const GemmTile* d_tiles, int ntiles, int tile_cols, const int* pair_src, const _Float16* src_hl,
                    const float* cscale, float* slab, int L, int* queues, hipStream_t st);

// tiles of at most width columns (default 64) over the classes' columns 3 * class_begin
void plan_tiles(const std::vector<int>& class_begin, const std::vector<int>& class_op, std::vector<GemmTile>& tiles,
                int width = 64);

}  // namespace fmm

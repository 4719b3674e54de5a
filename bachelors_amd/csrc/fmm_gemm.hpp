#pragma once
#include "fmm_common.hpp"

namespace fmm {

enum class Precision { F64 = 0, F32 = 1, BF16X3 = 2 };

struct GemmTile {
    int op;      // operator index
    int col0;    // first global column (3 * pair + component)
    int ncol;    // valid columns in this tile (<= 64)
};

// ops: operator tables (double / float / bf16 hi), ops_lo: bf16 lo parts (BF16X3 only).
// transpose: apply T^T (adjoint passes).
// src_exp / dst_exp: [node][3][Lp] in double (F64) or float (F32, BF16X3).
void class_gemm(Precision prec, const void* ops, const void* ops_lo, const GemmTile* d_tiles, int ntiles,
                const int* pair_src, const int* pair_dst, const void* src_exp, void* dst_exp, int Lp, bool atomic,
                hipStream_t st, bool transpose = false);

void plan_tiles(const std::vector<int>& class_begin, const std::vector<int>& class_op, std::vector<GemmTile>& tiles);

}  // namespace fmm

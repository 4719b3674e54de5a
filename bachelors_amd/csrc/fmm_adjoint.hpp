#pragma once
#include "fmm_gemm.hpp"
#include "fmm_lists.hpp"

namespace fmm {

struct PassView {
    int ntiles = 0;
    const GemmTile* tiles = nullptr;
    const int* src = nullptr;   // forward source of each pair
    const int* dst = nullptr;   // forward destination of each pair
    int width = 64;             // columns per tile the pass was planned with (class_gemm takes <= 64)
};

struct AdjointArgs {
    Precision prec;
    int p, Lp;
    const TreeDev* S;
    const TreeDev* T;
    ScaleTable sigS, sigT;
    const int* src_leaves;
    int n_src_leaves;
    const int* tgt_leaves;
    int n_tgt_leaves;
    const double* sxyz;   // slot order
    const double* sJ;     // slot order current directions
    const double* txyz;   // slot order
    const Lists* lists;
    int near_field;
    const void* up_ops;
    const void* up_ops_lo;
    const void* down_ops;
    const void* down_ops_lo;
    int n_up_ops, n_down_ops;
    std::vector<PassView> m2m_levels, l2l_levels;   // indexed by child level
    PassView m2l;                                   // src = far_t, dst = far_s (already transposed roles)
    double scale;
    int nt, ns;
};

struct AdjointState {
    DevBuf g, lam, mu, gq, Gexp;
    PairIndex m2l_ix;   // far pairs by source node
    DevBuf slab;        // transposed M2L result per (pair, component)
};

// dLdw (input order) = transposed FMM applied to dLdB (input order); device pointers.
void run_adjoint(AdjointState& st, const AdjointArgs& a, const double* dLdB, double* dLdw, hipStream_t stream);

}  // namespace fmm

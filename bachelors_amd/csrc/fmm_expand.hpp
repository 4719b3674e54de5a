#pragma once
#include "fmm_common.hpp"

namespace fmm {

void launch_charges(const int* slots, int n, const double* xyz, const double* w, const double* J, double* sxyz,
                    double* sq, double* sJ, hipStream_t st);
void launch_gather_xyz(const int* slots, int n, const double* xyz, double* out, hipStream_t st);
// F16X3: where P2M also writes its leaves' columns of the M2L's split B operand (see
// split_columns: coefficient scales dm, per-column scales, FP16 hi | lo per K step); hl null: none
struct F16Leaves {
    const float* dm = nullptr;
    int Kp = 0;
    float* cscale = nullptr;
    _Float16* hl = nullptr;
};
void launch_p2m(bool f64, const int* leaves, int nleaves, const TreeDev& S, const ScaleTable& sig, const double* sxyz,
                const double* sq, int p, int Lp, void* M, hipStream_t st, const F16Leaves& split = F16Leaves());
// clears the expansions (3 Lp elements of esize bytes per node) of every node launch_p2m does not
// write: subdivided nodes (the M2M adds into them) and leaves without points
void launch_zero_unwritten(const TreeDev& S, int Lp, size_t esize, void* M, hipStream_t st);
void launch_l2p(bool f64, const int* leaves, int nleaves, const TreeDev& T, const ScaleTable& sig, const double* txyz,
                const void* Lexp, int p, int Lp, double* Bfar, hipStream_t st);
// near-field sources in slot order relative to origin[3] (a device pointer); P0, P1: 4 x R per
// slot, P2: R per slot, R = double if f64 else float
void launch_pack_sources(bool f64, int n, const double* origin, const double* sxyz, const double* sq, void* P0,
                         void* P1, void* P2, hipStream_t st);
void launch_p2p(bool f64, const int* leaves, int nleaves, const TreeDev& T, const TreeDev& S, const int* close_begin,
                const int* close_end, const int* close_s, const void* P0, const void* P1, const void* P2,
                const double* txyz, double* Bnear, hipStream_t st);
// launch_p2p as a persistent grid of `grid` workgroups taking leaves from the counter *queue
// (reset on the stream first); a small register budget lets it share CUs with the M2L GEMM
void launch_p2p_queue(bool f64, const int* leaves, int nleaves, const TreeDev& T, const TreeDev& S,
                      const int* close_begin, const int* close_end, const int* close_s, const void* P0, const void* P1,
                      const void* P2, const double* txyz, double* Bnear, int* queue, int grid, hipStream_t st);
void launch_assemble(const int* slots, int n, const double* Bfar, const double* Bnear, double scale, double* B,
                     hipStream_t st);
// L2P / P2P of the vector potential A (the calcA counterparts of launch_l2p / launch_p2p)
void launch_l2p_A(bool f64, const int* leaves, int nleaves, const TreeDev& T, const ScaleTable& sig, const double* txyz,
                  const void* Lexp, int p, int Lp, double* Afar, hipStream_t st);
void launch_p2p_A(bool f64, const int* leaves, int nleaves, const TreeDev& T, const TreeDev& S, const int* close_begin,
                  const int* close_end, const int* close_s, const void* P0, const void* P1, const void* P2,
                  const double* txyz, double* Anear, hipStream_t st);
// Problem::BEM: slot-order strengths v = w (B x n) (3) and d = w (B . n) of BEMQuadrature sources;
// their multipoles (P2M of the gradient of R); the near field of their potential (FP64, unscaled)
void launch_bem_charges(const int* slots, int n, const double* xyz, const double* w, const double* Bv,
                        const double* nrm, double* sxyz, double* sv, double* sd, hipStream_t st);
void launch_p2m_bem(bool f64, const int* leaves, int nleaves, const TreeDev& S, const ScaleTable& sig,
                    const double* sxyz, const double* sv, const double* sd, int p, int Lp, void* M, hipStream_t st);
void launch_p2p_bem(const int* leaves, int nleaves, const TreeDev& T, const TreeDev& S, const int* close_begin,
                    const int* close_end, const int* close_s, const double* sxyz, const double* sv, const double* sd,
                    const double* txyz, double* out, hipStream_t st);
// FP64 direct sum of B (pot: A) at nt targets over the ns slot-order sources, scaled; the sources
// are cut into nsplit ranges whose partial sums (partial: nsplit * nt * 3 doubles) are added in
// range order.  direct_splits: the number of ranges that fills the GPU.
int direct_splits(int ns, int nt);
void launch_direct(bool pot, const double* sxyz, const double* sq, int ns, const double* txyz, int nt, double scale,
                   double* out, double* partial, int nsplit, hipStream_t st);

}  // namespace fmm

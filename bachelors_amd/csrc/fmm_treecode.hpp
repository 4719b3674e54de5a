#pragma once
#include "fmm_common.hpp"

namespace fmm {

// Treecode M2P over the source tree S (multipoles M: [node][3][Lp], double if f64 else float,
// scaled by sigma^-l): at n input-order points x, walking from the root; B (curl A) times
// scaleB and A times scaleA, either output may be null.
void launch_treecode_points(bool f64, const TreeDev& S, const ScaleTable& sig, const void* M, int p, int Lp,
                            const double* x, int n, double scaleB, double scaleA, double* B, double* A,
                            hipStream_t st);
// The same walk from every node of each useful target leaf's close list, at the leaf's
// slot-order targets txyz: the reference's close-list evaluation (no particle sums), unscaled;
// Bnear and Anear (slot order) may be null.
void launch_m2p_close(bool f64, const TreeDev& S, const ScaleTable& sig, const void* M, int p, int Lp,
                      const int* leaves, int nleaves, const TreeDev& T, const int* close_begin, const int* close_end,
                      const int* close_s, const double* txyz, double* Bnear, double* Anear, hipStream_t st);

}  // namespace fmm

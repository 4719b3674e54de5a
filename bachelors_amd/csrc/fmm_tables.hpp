#pragma once
#include "fmm_common.hpp"

namespace fmm {

enum OpKind { OPK_M2M = 0, OPK_M2L = 1, OPK_L2L = 2 };

// One translation operator: t = translation vector of the reference translator
// (M2M: child - parent, M2L: source - target, L2L: parent - child), sig_in / sig_out =
// expansion scales of the input / output nodes.
struct OpSpec {
    double t[3];
    double sig_in, sig_out;
    int kind;
    int m2l_full;   // M2L keeps l + lambda <= 2p (else the reference's l + lambda <= p)
};

// FP16 split tables (make_m2l_f16): the operator equilibrated as rs_i T_ij / dm_j -- dm_j the
// input coefficient's scale, rs_i the power of 2 that brings row i's maximum into [0.5, 1) --
// and split into hi + lo FP16 parts, [op][R rows][Kp columns] (R >= L, Kp >= L, zero padded);
// rinv[op][i] = 1 / rs_i restores the rows of T x.
struct F16Split {
    _Float16* hi = nullptr;
    _Float16* lo = nullptr;
    float* rinv = nullptr;
    const float* dm = nullptr;   // [>= L]
    int R = 0, Kp = 0;
};

// nops operators of Lp x Lp, row-major [out][in], zero padded, computed in FP64 and written as
// out_mode 0: double (d_out), 1: float (d_out), 2: bf16 hi (d_out) + bf16 lo (d_out_lo) of the
// float value.
void make_operators(const OpSpec* d_specs, int nops, int p, int Lp, int out_mode, void* d_out, __bf16* d_out_lo,
                    hipStream_t st);
// nops M2L operators (specs of kind OPK_M2L) as FP16 splits, computed in FP64; work: device
// scratch of m2l_f16_work_bytes(nops, p) (the harmonic tables)
size_t m2l_f16_work_bytes(int nops, int p);
void make_m2l_f16(const OpSpec* d_specs, int nops, int p, const F16Split& f16, void* work, hipStream_t st);
void make_m2l_specs(const int* class_begin, int nc, const int* far_t, const int* far_s, const double* sc,
                    const double* tc, const int* slev, const int* tlev, const ScaleTable& sigS, const ScaleTable& sigT,
                    int m2l_full, OpSpec* out, hipStream_t st);
// M2M of one tree level without operator tables: every parent in parents[0 .. nparents) adds its
// non-empty children's expansions translated by the complex formula (M: [node][3][Lp] scaled
// expansions, double if f64 else float)
void m2m_per_node(bool f64, const int* parents, int nparents, const TreeDev& S, const ScaleTable& sig, int p, int Lp,
                  void* M, hipStream_t st);
// the M2M operator specs per (child level, orientation), [(k - 1) * 8 + o], from the first non-empty
// node of each (the reference's choice, multipole_solver.cpp:472-534); d_nominal [(k - 1) * 8 + o][3]
// holds the offsets used where a level has no node of an orientation
void make_m2m_specs(const TreeDev& S, const ScaleTable& sig, const double* d_nominal, OpSpec* out, hipStream_t st);
void make_pair_specs(const int* src, const int* dst, int n, const double* csrc, const double* cdst, const int* lsrc,
                     const int* ldst, const ScaleTable& sig_src, const ScaleTable& sig_dst, int kind, int m2l_full,
                     OpSpec* out, hipStream_t st);

}  // namespace fmm

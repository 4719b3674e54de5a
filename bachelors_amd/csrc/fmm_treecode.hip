// Treecode evaluation of the source tree's multipole expansions (M2P): the reference's
// single-point MultipoleSolver::calcA / calcB(current, point) (multipole_solver.cpp:686-700 ->
// QuadratureOctreeNode::calcA / calcRot, quadrature_octree.cpp:171-258), and the same recursion
// started at the nodes of a target leaf's close list, which is how the reference's calcA / calcB
// (current) evaluate the close list (fast_multipole_solver.cpp:195-231).
//
// From a start node: a non-empty node whose box passes 4 R^2 < d^2 (R = |half sizes|, d = the
// distance from its centre) adds the M2P of its multipole expansion; a node that fails opens its
// 8 children; a leaf that fails adds nothing (the reference has no particle sum there).  The walk
// is stackless: the 8 children of a node are consecutive ids, so after a node the next one is
// its next sibling, or the next sibling of the nearest ancestor that has one.
//
// M2P: A_c(x) = sum_{l <= p, m} M_c(l, m) I_l^m(x - c) with the reference's real irregular
// harmonics I(l, 0) = (-1)^l Psi_l^0, I(l, +-m) = (-1)^l sqrt2 {Re, Im} Psi_l^m, where
// Psi_l^m = (l - m)! P_l^m(cos) e^{i m phi} / r^{l+1} follows the Cartesian recurrences
//   Psi_0^0 = 1/r, Psi_m^m = (2m - 1) (x + iy) Psi_{m-1}^{m-1} / r^2,
//   Psi_l^m = ((2l - 1) z Psi_{l-1}^m - ((l-1)^2 - m^2) Psi_{l-2}^m) / r^2.
// B = curl A from the exact gradient (no finite differences):
//   dz Psi_l^m = -Psi_{l+1}^m, (dx + i dy) Psi_l^m = -Psi_{l+1}^{m+1}, (dx - i dy) Psi_l^m = Psi_{l+1}^{m-1}
// with Psi_l^{-m} = (-1)^m conj(Psi_l^m).  With A_c = Re sum_{m >= 0} alpha_c(l, m) Psi_l^m, every
// Psi_n^k of degree <= p + 1 is used once, by the alpha of the terms it serves.  The multipoles
// are stored scaled (M_true = sigma^l M), so A = M2P(u) / sigma and grad A = grad_u / sigma^2 with
// u = (x - c) / sigma.  Everything runs in FP64.
#include "fmm_treecode.hpp"

namespace fmm {

constexpr double SQRT2_T = 1.41421356237309504880;

// alpha_c(l, m) = (-1)^l M_c(l, 0) (m = 0), (-1)^l sqrt2 (M_c(l, m) - i M_c(l, -m)) (m > 0)
template <class T>
__device__ __forceinline__ void alpha_at(const T* __restrict__ Mn, int Lp, int c, int l, int m, double& ar,
                                         double& ai) {
    const T* Mc = Mn + c * Lp;
    const double s = (l & 1) ? -1.0 : 1.0;
    const int k = l * l + l;
    if (m == 0) {
        ar = s * (double)Mc[k];
        ai = 0.0;
    } else {
        ar = s * SQRT2_T * (double)Mc[k + m];
        ai = -s * SQRT2_T * (double)Mc[k - m];
    }
}

// A (3) and grad A (G[c][axis]) of one node's expansion at the scaled offset u
template <class T>
__device__ void m2p_node(const T* __restrict__ Mn, int p, int Lp, double ux, double uy, double uz, double A[3],
                         double G[3][3]) {
    const double r2 = ux * ux + uy * uy + uz * uz, ir2 = 1.0 / r2;
    double dr = sqrt(ir2), di = 0.0;   // Psi_k^k
    for (int k = 0; k <= p + 1; k++) {
        if (k > 0) {
            const double f = (2.0 * k - 1) * ir2;
            const double nr = (dr * ux - di * uy) * f, ni = (dr * uy + di * ux) * f;
            dr = nr;
            di = ni;
        }
        double p1r = dr, p1i = di, p2r = 0.0, p2i = 0.0;
        for (int n = k; n <= p + 1; n++) {
            double vr = dr, vi = di;   // Psi_n^k
            if (n > k) {
                const double a = (2.0 * n - 1) * uz, b = (double)((n - 1) * (n - 1) - k * k);
                vr = (a * p1r - b * p2r) * ir2;
                vi = (a * p1i - b * p2i) * ir2;
                p2r = p1r; p2i = p1i; p1r = vr; p1i = vi;
            }
            for (int c = 0; c < 3; c++) {
                double ar, ai;
                if (n <= p) {   // A: alpha(n, k) Psi_n^k
                    alpha_at(Mn, Lp, c, n, k, ar, ai);
                    A[c] += ar * vr - ai * vi;
                }
                const int l = n - 1;   // gradient terms of degree l = n - 1 <= p
                if (l < 0 || l > p) continue;
                if (k <= l) {          // dz: -alpha(l, k) Psi_{l+1}^k
                    alpha_at(Mn, Lp, c, l, k, ar, ai);
                    G[c][2] -= ar * vr - ai * vi;
                }
                if (k + 1 <= l) {      // from Psi_{l+1}^{m-1}, m = k + 1: dx +1/2, dy +i/2
                    alpha_at(Mn, Lp, c, l, k + 1, ar, ai);
                    G[c][0] += 0.5 * (ar * vr - ai * vi);
                    G[c][1] -= 0.5 * (ar * vi + ai * vr);
                }
                if (k >= 1) {          // from Psi_{l+1}^{m+1}, m = k - 1: dx -1/2, dy +i/2
                    alpha_at(Mn, Lp, c, l, k - 1, ar, ai);
                    G[c][0] -= 0.5 * (ar * vr - ai * vi);
                    G[c][1] -= 0.5 * (ar * vi + ai * vr);
                    if (k == 1) {      // m = 0 also meets Psi_{l+1}^{-1} = -conj(Psi_{l+1}^1)
                        G[c][0] -= 0.5 * ar * vr;
                        G[c][1] -= 0.5 * ar * vi;
                    }
                }
            }
        }
    }
}

struct TreeView {
    const double* center;
    const int* first_child;
    const int* parent;
    const int* npts;
    const int* level;
    ScaleTable sig;
    double radius[MAX_LEVELS];
};

// the treecode walk from node start at point x (A and grad A accumulated, unscaled by current)
template <class T>
__device__ void walk(const TreeView& S, const T* __restrict__ M, int p, int Lp, int start, double x, double y,
                     double z, double A[3], double G[3][3]) {
    int n = start;
    for (;;) {
        bool open = false;
        const int fc = S.first_child[n];
        if (fc >= 0 || S.npts[n] > 0) {
            const double cx = S.center[3 * n], cy = S.center[3 * n + 1], cz = S.center[3 * n + 2];
            // Vector3::distanceSquared and 4 radius^2 < d^2 (quadrature_octree.cpp:177-181)
            const double dx = x - cx, dy = y - cy, dz = z - cz;
            const double d2 = dx * dx + dy * dy + dz * dz;
            const int lv = S.level[n];
            const double rad = S.radius[lv];
            if (4 * rad * rad < d2) {
                const double inv = 1.0 / S.sig.s[lv], i2 = inv * inv;
                double a[3] = {0, 0, 0}, g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                m2p_node(M + (size_t)n * 3 * Lp, p, Lp, dx * inv, dy * inv, dz * inv, a, g);
                for (int c = 0; c < 3; c++) {
                    A[c] += a[c] * inv;
                    for (int e = 0; e < 3; e++) G[c][e] += g[c][e] * i2;
                }
            } else if (fc >= 0) {
                open = true;
            }
        }
        if (open) {
            n = S.first_child[n];
            continue;
        }
        while (n != start && n - S.first_child[S.parent[n]] == 7) n = S.parent[n];
        if (n == start) break;
        n++;
    }
}

__device__ __forceinline__ void curl_of(const double G[3][3], double B[3]) {
    B[0] = G[2][1] - G[1][2];
    B[1] = G[0][2] - G[2][0];
    B[2] = G[1][0] - G[0][1];
}

// one thread per point, walk from the root (input-order points and outputs)
template <class T>
__global__ __launch_bounds__(64) void k_treecode_points(TreeView S, const T* __restrict__ M, int p, int Lp,
                                                        const double* __restrict__ x, int n, double scaleB,
                                                        double scaleA, double* __restrict__ B, double* __restrict__ A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a[3] = {0, 0, 0}, g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, b[3];
    walk(S, M, p, Lp, 0, x[3 * i], x[3 * i + 1], x[3 * i + 2], a, g);
    curl_of(g, b);
    for (int c = 0; c < 3; c++) {
        if (B) B[3 * i + c] = b[c] * scaleB;
        if (A) A[3 * i + c] = a[c] * scaleA;
    }
}

// one workgroup per useful target leaf, one thread per target: the walk from every node of the
// leaf's close list (slot-order targets and outputs, unscaled)
template <class T>
__global__ __launch_bounds__(64) void k_m2p_close(TreeView S, const T* __restrict__ M, int p, int Lp,
                                                  const int* __restrict__ leaves, const int* __restrict__ tbegin,
                                                  const int* __restrict__ tend, const int* __restrict__ close_begin,
                                                  const int* __restrict__ close_end, const int* __restrict__ close_s,
                                                  const double* __restrict__ txyz, double* __restrict__ Bnear,
                                                  double* __restrict__ Anear) {
    const int leaf = leaves[blockIdx.x];
    for (int k = tbegin[leaf] + threadIdx.x; k < tend[leaf]; k += blockDim.x) {
        double a[3] = {0, 0, 0}, g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, b[3];
        for (int ci = close_begin[leaf]; ci < close_end[leaf]; ci++)
            walk(S, M, p, Lp, close_s[ci], txyz[3 * k], txyz[3 * k + 1], txyz[3 * k + 2], a, g);
        curl_of(g, b);
        for (int c = 0; c < 3; c++) {
            if (Bnear) Bnear[3 * k + c] = b[c];
            if (Anear) Anear[3 * k + c] = a[c];
        }
    }
}

static TreeView view_of(const TreeDev& S, const ScaleTable& sig) {
    TreeView v;
    v.center = S.center.as<double>();
    v.first_child = S.first_child.as<int>();
    v.parent = S.parent.as<int>();
    v.npts = S.npts.as<int>();
    v.level = S.level.as<int>();
    v.sig = sig;
    for (int k = 0; k < MAX_LEVELS; k++) v.radius[k] = S.radius[k];
    return v;
}

void launch_treecode_points(bool f64, const TreeDev& S, const ScaleTable& sig, const void* M, int p, int Lp,
                            const double* x, int n, double scaleB, double scaleA, double* B, double* A,
                            hipStream_t st) {
    if (n <= 0 || S.num_nodes <= 0) return;
    const TreeView v = view_of(S, sig);
    if (f64)
        k_treecode_points<double><<<blocks(n, 64), 64, 0, st>>>(v, (const double*)M, p, Lp, x, n, scaleB, scaleA, B, A);
    else
        k_treecode_points<float><<<blocks(n, 64), 64, 0, st>>>(v, (const float*)M, p, Lp, x, n, scaleB, scaleA, B, A);
    FMM_HIP(hipGetLastError());
}

void launch_m2p_close(bool f64, const TreeDev& S, const ScaleTable& sig, const void* M, int p, int Lp,
                      const int* leaves, int nleaves, const TreeDev& T, const int* close_begin, const int* close_end,
                      const int* close_s, const double* txyz, double* Bnear, double* Anear, hipStream_t st) {
    if (nleaves <= 0) return;
    const TreeView v = view_of(S, sig);
    if (f64)
        k_m2p_close<double><<<nleaves, 64, 0, st>>>(v, (const double*)M, p, Lp, leaves, T.begin.as<int>(),
                                                    T.end.as<int>(), close_begin, close_end, close_s, txyz, Bnear,
                                                    Anear);
    else
        k_m2p_close<float><<<nleaves, 64, 0, st>>>(v, (const float*)M, p, Lp, leaves, T.begin.as<int>(),
                                                   T.end.as<int>(), close_begin, close_end, close_s, txyz, Bnear,
                                                   Anear);
    FMM_HIP(hipGetLastError());
}

}  // namespace fmm

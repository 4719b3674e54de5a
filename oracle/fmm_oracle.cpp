// fmm_oracle.cpp -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
//
// CPU FP64 restatement of the reference Biot-Savart fast multipole method
// (iamwhoamianamy/bachelors, FMM/FMMGPU, built with REAL_IS_DOUBLE).  It is the
// checker for the HIP library: only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it.  Nothing in bachelors_amd/ links it.
//
// Every function cites the reference file:line it restates.  The reference's
// numerical quirks are reproduced by default and can be switched off one by one
// (ORC_* flags below); "corrected" mode = all fixes on.
//
// Parity pins (what checks this restatement, see tests/test_oracle_*.py):
//   harmonics       : addition theorem 1/|x-y| = sum R(y) I(x); Cartesian
//                     recurrence route vs Legendre route; homogeneity;
//                     reference fixture matrices/regular_{0..7}.txt (complex M2M).
//   translators     : M2M/L2L exact identities, M2L->L2P convergence to 1/|x-y|.
//   trees           : brute-force recount of every node's box population,
//                     subdivision <=> count >= threshold, reference child order.
//   interaction maps: brute-force recursion on small trees + coverage invariant.
//   direct sum      : closed-form on-axis field of a circular current loop and of a
//                     thick solenoid; Telma fixture results/telma_results.txt.
//   FMM             : converges to the direct sum as p grows (corrected mode).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cfloat>
#include <vector>
#include <set>
#include <map>
#include <algorithm>
#include <chrono>
#include <complex>
#include <omp.h>

namespace orc {

// ---- constants: FMM/FMMGPU/math.hpp:12-16 (truncated literals kept) ------------
static const double PI = 3.14159265359;
static const double MU0 = 1.2566370614e-6;
static const double FD_EPS = 1e-3;             // calculation_point_octree.cpp:217
static const double R_SQRT_2 = 0.70710678118;  // math.hpp:16

enum Flags {
    SECTORAL_FIX = 1,   // addComplex loops m <= p   (reference m < p, harmonics.cpp:130)
    FACT_FIX = 2,       // exact factorials          (reference casts to size_t, factorials.hpp:12)
    ANALYTIC_CURL = 4,  // L2P by analytic gradient  (reference FD, calculation_point_octree.cpp:217-264)
    P2P = 8,            // near field by direct sum  (reference treecode M2P, quadrature_octree.cpp:196-258)
    M2L_FULL = 16,      // M2L keeps l+lambda <= 2p  (reference l+lambda <= p, multipole_translator.cpp:165)
    FLOAT_EPS = 32,     // harmonics add FLT_EPSILON to r, as the default float build does (harmonics.cpp:28)
};

typedef std::complex<double> cd;

inline int lm(int l, int m) { return l * l + l + m; }   // harmonic_series.hpp:75-86
inline int nlen(int p) { return (p + 1) * (p + 1); }
inline double neg1pow(int k) { return (k % 2 == 0) ? 1.0 : -1.0; }

// ---- factorials: factorials.cpp:3-17 --------------------------------------------
struct Factorials {
    double f[200];
    Factorials() { f[0] = 1; for (int i = 1; i < 200; i++) f[i] = f[i - 1] * i; }
    // The reference returns size_t (factorials.hpp:12).  Exact for i <= 20; above that
    // the conversion overflows (undefined behaviour) and only the FACT_FIX path is valid.
    double get(int i, int flags) const {
        if (flags & FACT_FIX) return f[i];
        if (i > 20) return NAN;
        return (double)(uint64_t)f[i];
    }
};
static const Factorials FACT;

// ---- solid harmonics, Legendre route: harmonics.cpp:20-56, 68-140 ---------------
// out has (order+1)^2 entries, index l^2+l+m.
static void solid_harmonics(int order, double vx, double vy, double vz, bool regular,
                            int flags, double* Y) {
    const int n = nlen(order);
    double r = std::sqrt(vx * vx + vy * vy + vz * vz) + ((flags & FLOAT_EPS) ? (double)FLT_EPSILON : DBL_EPSILON);   // harmonics.cpp:24-28
    double x = vx / r, y = vy / r, z = vz / r;                           // point /= r
    for (int i = 0; i < n; i++) Y[i] = 0;
    // fillWithLegendrePolynomials (harmonics.cpp:77-86)
    Y[lm(0, 0)] = 1;
    if (order >= 1) Y[lm(1, 0)] = z;
    for (int l = 2; l <= order; l++)
        Y[lm(l, 0)] = ((2 * l - 1) * z * Y[lm(l - 1, 0)] - (l - 1) * Y[lm(l - 2, 0)]) / l;
    // fillWithLegendrePolynomialDerivatives (harmonics.cpp:88-106)
    Y[lm(0, 0)] = 1;
    if (order >= 1) { Y[lm(1, -1)] = 1; Y[lm(1, 1)] = 1; }
    for (int l = 2; l <= order; l++) {
        for (int m = 1; m < l; m++)
            Y[lm(l, m)] = z * Y[lm(l - 1, m)] + (l + m - 1) * Y[lm(l - 1, m - 1)];
        Y[lm(l, l)] = (2 * l - 1) * Y[lm(l - 1, l - 1)];
    }
    // mirrorLegendrePolynomialDerivatives (harmonics.cpp:108-117)
    for (int l = 2; l <= order; l++)
        for (int m = -l; m <= l; m++) Y[lm(l, m)] = Y[lm(l, -m)];
    // addComplex (harmonics.cpp:125-140); the reference stops at m < order.
    cd e(std::sqrt(2.0), 0.0), mult(x, y);
    const int mmax = (flags & SECTORAL_FIX) ? order : order - 1;
    for (int m = 1; m <= mmax; m++) {
        e = cd(e.real() * mult.real() - e.imag() * mult.imag(),
               e.real() * mult.imag() + e.imag() * mult.real());
        for (int l = m; l <= order; l++) {
            Y[lm(l, m)] *= e.real();
            Y[lm(l, -m)] *= e.imag();
        }
    }
    // radial part and factorials (harmonics.cpp:32-53)
    double mul = regular ? r : 1.0 / r;
    double curr = regular ? 1.0 : mul;
    for (int l = 0; l <= order; l++) {
        for (int m = 0; m <= l; m++)
            Y[lm(l, -m)] *= curr * (regular ? 1.0 / FACT.get(l + m, flags) : FACT.get(l - m, flags));
        for (int m = 1; m <= l; m++)
            Y[lm(l, m)] *= curr * (regular ? 1.0 / FACT.get(l + m, flags) : FACT.get(l - m, flags));
        curr *= -mul;
    }
}

// ---- solid harmonics, Cartesian recurrence route (independent derivation) -------
// Phi_l^m = r^l P_l^m(cos t) e^{im phi}/(l+m)!,  Psi_l^m = (l-m)! P_l^m e^{im phi}/r^{l+1}
// (m >= 0, no Condon-Shortley phase).  Reference real basis: R(l,0) = (-1)^l Phi_l^0,
// R(l,m) = (-1)^l sqrt2 Re Phi_l^m, R(l,-m) = (-1)^l sqrt2 Im Phi_l^m; same for I with Psi.
static void phi_table(int order, double x, double y, double z, std::vector<cd>& P) {
    P.assign((order + 1) * (order + 1), cd(0, 0));
    auto at = [&](int l, int m) -> cd& { return P[l * (order + 1) + m]; };
    double r2 = x * x + y * y + z * z;
    at(0, 0) = 1;
    for (int m = 0; m <= order; m++) {
        if (m > 0) at(m, m) = at(m - 1, m - 1) * cd(x, y) / (2.0 * m);
        if (m + 1 <= order) at(m + 1, m) = z * at(m, m);
        for (int l = m + 2; l <= order; l++)
            at(l, m) = ((2.0 * l - 1) * z * at(l - 1, m) - r2 * at(l - 2, m)) / double(l * l - m * m);
    }
}
static void psi_table(int order, double x, double y, double z, std::vector<cd>& P) {
    P.assign((order + 1) * (order + 1), cd(0, 0));
    auto at = [&](int l, int m) -> cd& { return P[l * (order + 1) + m]; };
    double r2 = x * x + y * y + z * z;
    at(0, 0) = 1.0 / std::sqrt(r2);
    for (int m = 0; m <= order; m++) {
        if (m > 0) at(m, m) = at(m - 1, m - 1) * cd(x, y) * ((2.0 * m - 1) / r2);
        if (m + 1 <= order) at(m + 1, m) = (2.0 * m + 1) * z * at(m, m) / r2;
        for (int l = m + 2; l <= order; l++)
            at(l, m) = ((2.0 * l - 1) * z * at(l - 1, m) - double((l - 1) * (l - 1) - m * m) * at(l - 2, m)) / r2;
    }
}
static void complex_table_to_real(int order, const std::vector<cd>& P, double* out) {
    const double s2 = std::sqrt(2.0);
    for (int l = 0; l <= order; l++) {
        double sg = neg1pow(l);
        out[lm(l, 0)] = sg * P[l * (order + 1)].real();
        for (int m = 1; m <= l; m++) {
            out[lm(l, m)] = sg * s2 * P[l * (order + 1) + m].real();
            out[lm(l, -m)] = sg * s2 * P[l * (order + 1) + m].imag();
        }
    }
}
static void regular_cart(int order, double x, double y, double z, double* out) {
    std::vector<cd> P; phi_table(order, x, y, z, P); complex_table_to_real(order, P, out);
}
static void irregular_cart(int order, double x, double y, double z, double* out) {
    std::vector<cd> P; psi_table(order, x, y, z, P); complex_table_to_real(order, P, out);
}
// Gradient of the real regular harmonics: dz Phi_l^m = Phi_{l-1}^m,
// (dx - i dy) Phi_l^m = Phi_{l-1}^{m-1}, (dx + i dy) Phi_l^m = -Phi_{l-1}^{m+1},
// with Phi_l^{-m} = (-1)^m conj(Phi_l^m).  out: (order+1)^2 x 3 (d/dx, d/dy, d/dz).
static void regular_grad_cart(int order, double x, double y, double z, double* out) {
    std::vector<cd> P;
    phi_table(order, x, y, z, P);
    auto ph = [&](int l, int m) -> cd {
        if (l < 0 || std::abs(m) > l) return cd(0, 0);
        if (m >= 0) return P[l * (order + 1) + m];
        return neg1pow(-m) * std::conj(P[l * (order + 1) - m]);
    };
    const double s2 = std::sqrt(2.0);
    for (int l = 0; l <= order; l++) {
        double sg = neg1pow(l);
        for (int m = 0; m <= l; m++) {
            cd dm = ph(l - 1, m - 1), dp = -ph(l - 1, m + 1);   // D- Phi, D+ Phi
            cd gx = 0.5 * (dp + dm);
            cd gy = (dp - dm) / cd(0, 2);
            cd gz = ph(l - 1, m);
            if (m == 0) {
                out[3 * lm(l, 0) + 0] = sg * gx.real();
                out[3 * lm(l, 0) + 1] = sg * gy.real();
                out[3 * lm(l, 0) + 2] = sg * gz.real();
            } else {
                out[3 * lm(l, m) + 0] = sg * s2 * gx.real();
                out[3 * lm(l, m) + 1] = sg * s2 * gy.real();
                out[3 * lm(l, m) + 2] = sg * s2 * gz.real();
                out[3 * lm(l, -m) + 0] = sg * s2 * gx.imag();
                out[3 * lm(l, -m) + 1] = sg * s2 * gy.imag();
                out[3 * lm(l, -m) + 2] = sg * s2 * gz.imag();
            }
        }
    }
}

// ---- real <-> complex: harmonics.cpp:281-320 ------------------------------------
static void real_to_complex(int order, const double* R, cd* C) {
    const double c = R_SQRT_2;
    for (int l = 0; l <= order; l++) {
        C[lm(l, 0)] = cd(R[lm(l, 0)], 0);
        for (int m = 1; m <= l; m++) {
            C[lm(l, m)] = cd(R[lm(l, m)] * c, R[lm(l, -m)] * c);
            C[lm(l, -m)] = cd(R[lm(l, m)] * c, -R[lm(l, -m)] * c);
        }
    }
}
static void complex_to_real(int order, const cd* C, double* R) {
    const double c = 1.0 / std::sqrt(2.0);
    for (int l = 0; l <= order; l++) {
        R[lm(l, 0)] = C[lm(l, 0)].real();
        for (int m = 1; m <= l; m++) {
            R[lm(l, m)] = c * (C[lm(l, m)] + C[lm(l, -m)]).real();
            R[lm(l, -m)] = c * (C[lm(l, m)] - C[lm(l, -m)]).imag();
        }
    }
}

// ---- translation factors: multipole_translator.cpp:143-146, 195-199, 251-255 ----
inline double m2m_factor(int m, int mu) {
    return neg1pow(std::abs((std::abs(m) - std::abs(mu) - std::abs(m - mu)) / 2));
}
inline double m2l_factor(int m, int mu, int lambda) {
    return neg1pow(std::abs((std::abs(m - mu) - std::abs(m) - std::abs(mu)) / 2)) * neg1pow(lambda);
}
inline double l2l_factor(int m, int mu, int lambda, int l) {
    return neg1pow(std::abs((std::abs(mu) - std::abs(m - mu) - std::abs(m)) / 2)) * neg1pow(lambda + l);
}

// ---- complex translators --------------------------------------------------------
// M2M, multipole_translator.cpp:3-32: res(l,m) += R(lambda,mu) b(l-lambda, m-mu) f(m,mu)
static void translate_multipole_c(int p, const cd* R, const cd* b, cd* res) {
    for (int i = 0; i < nlen(p); i++) res[i] = 0;
    for (int l = 0; l <= p; l++)
        for (int m = -l; m <= l; m++)
            for (int lambda = 0; lambda <= l; lambda++) {
                int dl = l - lambda;
                for (int mu = -lambda; mu <= lambda; mu++) {
                    int dm = m - mu;
                    if (dm >= -dl && dm <= dl) res[lm(l, m)] += R[lm(lambda, mu)] * b[lm(dl, dm)] * m2m_factor(m, mu);
                }
            }
}
// M2L, multipole_translator.cpp:148-176: res(l,m) += a(lambda,mu) I(l+lambda, m-mu) g
// I has order pI (p, or 2p with M2L_FULL); terms with l+lambda > pI are dropped.
static void multipole_to_local_c(int p, int pI, const cd* a, const cd* I, cd* res) {
    for (int i = 0; i < nlen(p); i++) res[i] = 0;
    for (int l = 0; l <= p; l++)
        for (int lambda = 0; lambda <= p; lambda++) {
            int dl = l + lambda;
            if (dl > pI) continue;
            for (int m = -l; m <= l; m++)
                for (int mu = -lambda; mu <= lambda; mu++) {
                    int dm = m - mu;
                    if (dm >= -dl && dm <= dl) res[lm(l, m)] += a[lm(lambda, mu)] * I[lm(dl, dm)] * m2l_factor(m, mu, lambda);
                }
        }
}
// L2L, multipole_translator.cpp:201-232: res(l,m) += a(lambda,mu) R(lambda-l, m-mu) h
static void translate_local_c(int p, const cd* a, const cd* R, cd* res) {
    for (int i = 0; i < nlen(p); i++) res[i] = 0;
    for (int l = 0; l <= p; l++)
        for (int lambda = l; lambda <= p; lambda++)
            for (int m = -l; m <= l; m++)
                for (int mu = -lambda; mu <= lambda; mu++) {
                    int dl = lambda - l, dm = m - mu;
                    if (dm >= -dl && dm <= dl) res[lm(l, m)] += a[lm(lambda, mu)] * R[lm(dl, dm)] * l2l_factor(m, mu, lambda, l);
                }
}

// "WithComplex" wrappers for one real component (multipole_translator.cpp:111-125, 178-193, 234-249)
enum OpKind { OP_M2M = 0, OP_M2L = 1, OP_L2L = 2 };
struct Translator {
    int p, pI, flags, kind;
    std::vector<double> harm;   // real harmonics of the translation vector
    std::vector<cd> harmc;
    Translator(int kind_, int p_, const double t[3], int flags_) : p(p_), flags(flags_), kind(kind_) {
        pI = (kind == OP_M2L && (flags & M2L_FULL)) ? 2 * p : p;
        harm.resize(nlen(pI));
        harmc.resize(nlen(pI));
        solid_harmonics(pI, t[0], t[1], t[2], kind != OP_M2L, flags, harm.data());
        real_to_complex(pI, harm.data(), harmc.data());
    }
    void apply(const double* in, double* out) const {
        std::vector<cd> a(nlen(p)), r(nlen(p));
        real_to_complex(p, in, a.data());
        if (kind == OP_M2M) translate_multipole_c(p, harmc.data(), a.data(), r.data());
        else if (kind == OP_M2L) multipole_to_local_c(p, pI, a.data(), harmc.data(), r.data());
        else translate_local_c(p, a.data(), harmc.data(), r.data());
        complex_to_real(p, r.data(), out);
    }
};

// ---- M2M operator of the Matrices path: multipole_solver.cpp:472-534, 569-598 ----
// Real operator T (L x L, row-major, x_out = T x_in) = transpose of Re(C_r2c M_o C_c2r).
static void m2m_matrix_operator(int p, const double t[3], int flags, double* T) {
    const int L = nlen(p);
    std::vector<double> Rr(L);
    std::vector<cd> R(L);
    solid_harmonics(p, t[0], t[1], t[2], true, flags, Rr.data());
    real_to_complex(p, Rr.data(), R.data());
    std::vector<cd> M((size_t)L * L), A((size_t)L * L), B((size_t)L * L), t1((size_t)L * L);
    // formMatrixFromRegularHarmonicsForM2MAsVectors: M[in][out]
    for (int l = 0; l <= p; l++)
        for (int m = -l; m <= l; m++)
            for (int lambda = 0; lambda <= l; lambda++) {
                int dl = l - lambda;
                for (int mu = -lambda; mu <= lambda; mu++) {
                    int dm = m - mu;
                    if (-dl <= dm && dm <= dl) M[(size_t)lm(dl, dm) * L + lm(l, m)] = R[lm(lambda, mu)] * m2m_factor(m, mu);
                }
            }
    // calcRealToComplexTransitionMatrix1D / calcComplexToRealTransitionMatrix1D (harmonics.cpp:191-244)
    for (int l = 0; l <= p; l++) {
        for (int m = -l; m <= l; m++) {
            double rp = (m != 0) ? R_SQRT_2 : 1.0;
            double ip = (m > 0 ? 1 : (m < 0 ? -1 : 0)) * R_SQRT_2;
            int cur = lm(l, m), nr = lm(l, std::abs(m)), ni = lm(l, -std::abs(m));
            A[(size_t)nr * L + cur] += cd(rp, 0);
            A[(size_t)ni * L + cur] += cd(0, ip);
        }
        B[(size_t)lm(l, 0) * L + lm(l, 0)] = cd(1, 0);
        for (int m = 1; m <= l; m++) {
            B[(size_t)lm(l, m) * L + lm(l, m)] = cd(R_SQRT_2, 0);
            B[(size_t)lm(l, -m) * L + lm(l, m)] = cd(R_SQRT_2, 0);
            B[(size_t)lm(l, m) * L + lm(l, -m)] = cd(0, -R_SQRT_2);
            B[(size_t)lm(l, -m) * L + lm(l, -m)] = cd(0, R_SQRT_2);
        }
    }
    for (int i = 0; i < L; i++)
        for (int k = 0; k < L; k++) {
            cd a = A[(size_t)i * L + k];
            if (a == cd(0, 0)) continue;
            for (int j = 0; j < L; j++) t1[(size_t)i * L + j] += a * M[(size_t)k * L + j];
        }
    for (int i = 0; i < L; i++)
        for (int j = 0; j < L; j++) {
            cd s = 0;
            for (int k = 0; k < L; k++) s += t1[(size_t)i * L + k] * B[(size_t)k * L + j];
            T[(size_t)j * L + i] = s.real();   // row-major [in][out] -> store as [out][in]
        }
}

// ---- trees: quadrature_octree.cpp:23-97, calculation_point_octree.cpp:26-97, box.cpp:13-24
struct Node {
    double c[3], h[3];
    int parent = -1, level = 0;
    int first_child = -1;        // 8 consecutive children when subdivided
    std::vector<int> pts;        // arrivals kept while a leaf (insertion order)
};
struct Tree {
    std::vector<Node> nodes;     // insertion-order storage
    int thr;                     // leaf holds at most thr-1 points
    int max_depth;
    const double* P;             // xyz, 3 per point
    // canonical BFS numbering and DFS slot order
    std::vector<int> bfs;        // bfs id -> storage id
    std::vector<int> sid2bfs;
    std::vector<int> slots;      // DFS leaf order of point indices
    std::vector<int> begin, end; // per bfs node
    int dropped = 0;
};

static inline bool in_range(double v, double c, double h) { return c - h <= v && v < c + h; }   // box.cpp:13-18
static inline bool box_contains(const Node& n, const double* p) {
    return in_range(p[0], n.c[0], n.h[0]) && in_range(p[1], n.c[1], n.h[1]) && in_range(p[2], n.c[2], n.h[2]);
}

// child order and float-rounded centres: quadrature_octree.cpp:74-97
static const int CHILD_SIGN[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                     {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
static void subdivide(Tree& T, int ni) {
    float x = (float)T.nodes[ni].c[0], y = (float)T.nodes[ni].c[1], z = (float)T.nodes[ni].c[2];
    float w = (float)T.nodes[ni].h[0], h = (float)T.nodes[ni].h[1], d = (float)T.nodes[ni].h[2];
    int first = (int)T.nodes.size();
    for (int o = 0; o < 8; o++) {
        Node c;
        float cx = CHILD_SIGN[o][0] < 0 ? x - w / 2 : x + w / 2;
        float cy = CHILD_SIGN[o][1] < 0 ? y - h / 2 : y + h / 2;
        float cz = CHILD_SIGN[o][2] < 0 ? z - d / 2 : z + d / 2;
        c.c[0] = cx; c.c[1] = cy; c.c[2] = cz;
        c.h[0] = w / 2; c.h[1] = h / 2; c.h[2] = d / 2;
        c.parent = ni;
        c.level = T.nodes[ni].level + 1;
        T.nodes.push_back(c);
    }
    T.nodes[ni].first_child = first;
}
static void insert(Tree& T, int ni, int pi) {
    if (!box_contains(T.nodes[ni], T.P + 3 * (size_t)pi)) return;
    Node* n = &T.nodes[ni];
    if ((int)n->pts.size() + 1 < T.thr || n->level >= T.max_depth) {
        if (n->first_child < 0) n->pts.push_back(pi);
        else for (int o = 0; o < 8; o++) insert(T, T.nodes[ni].first_child + o, pi);
    } else {
        subdivide(T, ni);
        std::vector<int> pts;
        pts.swap(T.nodes[ni].pts);
        pts.push_back(pi);
        for (int q : pts)
            for (int o = 0; o < 8; o++) insert(T, T.nodes[ni].first_child + o, q);
    }
}
static void finalize_tree(Tree& T) {
    const int n = (int)T.nodes.size();
    T.bfs.clear();
    T.sid2bfs.assign(n, -1);
    T.bfs.push_back(0);
    for (size_t i = 0; i < T.bfs.size(); i++) {
        int s = T.bfs[i];
        if (T.nodes[s].first_child >= 0)
            for (int o = 0; o < 8; o++) T.bfs.push_back(T.nodes[s].first_child + o);
    }
    for (int i = 0; i < n; i++) T.sid2bfs[T.bfs[i]] = i;
    T.begin.assign(n, 0);
    T.end.assign(n, 0);
    T.slots.clear();
    // DFS in child order
    std::vector<std::pair<int, int>> st;   // (storage id, state)
    st.push_back({0, 0});
    while (!st.empty()) {
        auto [s, state] = st.back();
        st.pop_back();
        int b = T.sid2bfs[s];
        if (state == 0) {
            T.begin[b] = (int)T.slots.size();
            if (T.nodes[s].first_child < 0) {
                for (int q : T.nodes[s].pts) T.slots.push_back(q);
                T.end[b] = (int)T.slots.size();
            } else {
                st.push_back({s, 1});
                for (int o = 7; o >= 0; o--) st.push_back({T.nodes[s].first_child + o, 0});
            }
        } else {
            T.end[b] = (int)T.slots.size();
        }
    }
}
// source tree root Box((0,0,0),(3,3,3)): multipole_solver.cpp:662-663
// target tree root math::getBoundingBox: math.cpp:39-57
static void bounding_box(size_t n, const double* P, double c[3], double h[3]) {
    for (int a = 0; a < 3; a++) {
        double mx = -1e+16, mn = 1e+16;
        for (size_t i = 0; i < n; i++) {
            double v = P[3 * i + a];
            if (v > mx) mx = v;
            if (v < mn) mn = v;
        }
        double hd = (mx - mn) / 2 + 1e-5;
        c[a] = mn + hd;
        h[a] = hd * 1.1;
    }
}
static Tree* build_tree(size_t n, const double* P, const double root_c[3], const double root_h[3],
                        int thr, int max_depth) {
    Tree* T = new Tree();
    T->thr = thr;
    T->max_depth = max_depth;
    T->P = P;
    Node root;
    for (int a = 0; a < 3; a++) { root.c[a] = root_c[a]; root.h[a] = root_h[a]; }
    T->nodes.reserve(n / 4 + 16);
    T->nodes.push_back(root);
    for (size_t i = 0; i < n; i++) insert(*T, 0, (int)i);
    finalize_tree(*T);
    return T;
}
inline bool is_sub(const Tree& T, int s) { return T.nodes[s].first_child >= 0; }
inline bool is_useful(const Tree& T, int s) { return !T.nodes[s].pts.empty(); }
inline double box_radius(const Node& n) {   // box.cpp:41-46
    return std::sqrt(n.h[0] * n.h[0] + n.h[1] * n.h[1] + n.h[2] * n.h[2]);
}

// ---- interaction maps: fast_multipole_solver.cpp:105-184 ------------------------
struct Lists {
    std::set<std::pair<int, int>> far, close;   // (target storage id, source storage id)
};
static bool far_enough(const Tree& S, int s, const Tree& T, int t, double mac_scale) {
    double r1 = box_radius(S.nodes[s]), r2 = box_radius(T.nodes[t]);
    double mn = std::min(r1, r2), mx = std::max(r1, r2);
    double md = 2 * mx + mn;
    md *= mac_scale;
    double dx = S.nodes[s].c[0] - T.nodes[t].c[0];
    double dy = S.nodes[s].c[1] - T.nodes[t].c[1];
    double dz = S.nodes[s].c[2] - T.nodes[t].c[2];
    double d2 = dx * dx + dy * dy + dz * dz;
    return md * md < d2;
}
// tie_s: a planted mistake for the tests only (descend the source side on R_t == R_s, i.e.
// "<=" in the rule below); the restatement is tie_s = false
static void form_maps(const Tree& S, int s, const Tree& T, int t, double mac, Lists& out, bool tie_s = false) {
    if (!((is_sub(S, s) || is_useful(S, s)) && (is_sub(T, t) || is_useful(T, t)))) return;
    if (far_enough(S, s, T, t, mac)) {
        out.far.insert({t, s});
    } else if (is_useful(T, t)) {
        bool close_enough = true;
        if (is_sub(S, s))
            for (int o = 0; o < 8; o++)
                if (far_enough(S, S.nodes[s].first_child + o, T, t, mac)) { close_enough = false; break; }
        if (close_enough) out.close.insert({t, s});
        else for (int o = 0; o < 8; o++) form_maps(S, S.nodes[s].first_child + o, T, t, mac, out, tie_s);
    } else {
        const double rt = box_radius(T.nodes[t]), rs = box_radius(S.nodes[s]);
        if (((tie_s ? rt <= rs : rt < rs) || is_useful(T, t)) && !is_useful(S, s)) {
            for (int o = 0; o < 8; o++) form_maps(S, S.nodes[s].first_child + o, T, t, mac, out, tie_s);
        } else {
            for (int o = 0; o < 8; o++) form_maps(S, s, T, T.nodes[t].first_child + o, mac, out, tie_s);
        }
    }
}

// ---- direct Biot-Savart: integration.cpp:101-120 --------------------------------
inline void phi_hat(const double* y, double J[3]) {   // perp().normalized(), vector3.cu:154-167,201-204
    double px = -y[1], py = y[0];
    double l = std::sqrt(px * px + py * py);
    if (l != 0) { J[0] = px / l; J[1] = py / l; } else { J[0] = 0; J[1] = 0; }
    J[2] = 0;
}
inline void bs_kernel(const double* x, const double* y, const double q[3], double acc[3]) {
    double d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 == 0) return;   // coincident points contribute nothing (the reference divides by 0)
    double r = std::sqrt(r2);
    double r3 = r * r * r;
    acc[0] += (q[1] * d[2] - q[2] * d[1]) / r3;
    acc[1] += (q[2] * d[0] - q[0] * d[2]) / r3;
    acc[2] += (q[0] * d[1] - q[1] * d[0]) / r3;
}

// ---- the FMM: multipole_solver.cpp, fast_multipole_solver.cpp -------------------
struct Config {
    int p = 10;
    int cap_s = 32, cap_t = 32;
    int flags = 0;
    int target_root_mode = 0;        // 0: reference bounding box, 1: explicit box
    double target_root_c[3] = {0, 0, 0}, target_root_h[3] = {3, 3, 3};
    double source_root_c[3] = {0, 0, 0}, source_root_h[3] = {3, 3, 3};
    double mac_scale = 1.0;
    int max_depth = 21;
    double current = 1.0;
    double fd_eps = FD_EPS;          // FD curl step (reference 1e-3: calculation_point_octree.cpp:217,
                                     // quadrature_octree.cpp:201)
    int problem = 0;                 // 0 BioSavartLaplace, 1 BEM (multipole_solver_enums.hpp:26-30)
};
struct Run {
    Config cfg;
    Tree* S = nullptr;
    Tree* T = nullptr;
    Lists lists;
    std::vector<double> q;           // per source: w*J (3)
    std::vector<double> bem_d, bem_v;   // BEM per source: w (B.n) and w (B x n) (3)
    std::vector<std::vector<double>> M, Lc;   // per storage node: 3*L, component-major
    std::vector<double> B, Bfar, Bnear;       // per target slot (3 each)
    std::vector<double> A, Afar, Anear;       // vector potential per target slot (calcA(current))
    double t_tree = 0, t_p2m = 0, t_m2m = 0, t_lists = 0, t_m2l = 0, t_l2l = 0, t_eval = 0;
    ~Run() { delete S; delete T; }
};
typedef std::chrono::steady_clock clk;
static double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// BEM P2M: integration.cpp:295-353, M_c -= w (gradR_c (B.n) + (gradR x (B x n))_c) with gradR
// by central differences of step eps (reference) or, with ANALYTIC_CURL, the exact gradient
static void p2m_leaf_bem(const Run& R, const Tree& S, int s, std::vector<double>& out) {
    const int p = R.cfg.p, L = nlen(p);
    out.assign(3 * L, 0.0);
    std::vector<double> h1(L), h2(L), g(3 * L), ga(3 * L);
    const double e = R.cfg.fd_eps;
    for (int qi : S.nodes[s].pts) {
        const double* y = S.P + 3 * (size_t)qi;
        const double dx = y[0] - S.nodes[s].c[0], dy = y[1] - S.nodes[s].c[1], dz = y[2] - S.nodes[s].c[2];
        if (R.cfg.flags & ANALYTIC_CURL) {
            regular_grad_cart(p, dx, dy, dz, ga.data());
            for (int i = 0; i < L; i++)
                for (int a = 0; a < 3; a++) g[a * L + i] = ga[3 * i + a];
        } else {
            const double off[3][3] = {{e, 0, 0}, {0, e, 0}, {0, 0, e}};
            for (int a = 0; a < 3; a++) {
                solid_harmonics(p, dx + off[a][0], dy + off[a][1], dz + off[a][2], true, R.cfg.flags, h1.data());
                solid_harmonics(p, dx - off[a][0], dy - off[a][1], dz - off[a][2], true, R.cfg.flags, h2.data());
                for (int i = 0; i < L; i++) g[a * L + i] = (h1[i] - h2[i]) / (2 * e);
            }
        }
        const double d = R.bem_d[qi], *v = R.bem_v.data() + 3 * (size_t)qi;
        for (int i = 0; i < L; i++) {
            const double gx = g[i], gy = g[L + i], gz = g[2 * L + i];
            out[i] -= gx * d + (gy * v[2] - gz * v[1]);
            out[L + i] -= gy * d + (gz * v[0] - gx * v[2]);
            out[2 * L + i] -= gz * d + (gx * v[1] - gy * v[0]);
        }
    }
}
// the BEM integrand of one source at x (integration.cpp:191-202), times its weight:
// w ((y - x)(B.n) + (y - x) x (B x n)) / |y - x|^3
inline void bem_kernel(const double* x, const double* y, double d, const double* v, double acc[3]) {
    const double q[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
    const double r2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    if (r2 == 0) return;
    const double r3 = r2 * std::sqrt(r2);
    acc[0] += (q[0] * d + (q[1] * v[2] - q[2] * v[1])) / r3;
    acc[1] += (q[1] * d + (q[2] * v[0] - q[0] * v[2])) / r3;
    acc[2] += (q[2] * d + (q[0] * v[1] - q[1] * v[0])) / r3;
}

// P2M: integration.cpp:137-161 (one component-major expansion per leaf)
static void p2m_leaf(const Run& R, const Tree& S, int s, std::vector<double>& out) {
    if (R.cfg.problem == 1) { p2m_leaf_bem(R, S, s, out); return; }
    const int p = R.cfg.p, L = nlen(p);
    out.assign(3 * L, 0.0);
    std::vector<double> Rh(L);
    for (int qi : S.nodes[s].pts) {
        const double* y = S.P + 3 * (size_t)qi;
        solid_harmonics(p, y[0] - S.nodes[s].c[0], y[1] - S.nodes[s].c[1], y[2] - S.nodes[s].c[2], true,
                        R.cfg.flags, Rh.data());
        for (int i = 0; i < L; i++)
            for (int c = 0; c < 3; c++) out[c * L + i] += R.q[3 * (size_t)qi + c] * Rh[i];
    }
}

// treecode M2P with FD curl: quadrature_octree.cpp:196-258 (close list, reference behaviour)
static void m2p_rot(const Run& R, const Tree& S, int s, const double* x, double out[3]) {
    const int p = R.cfg.p, L = nlen(p);
    const Node& n = S.nodes[s];
    double dx = x[0] - n.c[0], dy = x[1] - n.c[1], dz = x[2] - n.c[2];
    double d2 = dx * dx + dy * dy + dz * dz;
    double rad = box_radius(n);
    if (4 * rad * rad < d2) {
        if (R.M[s].empty()) return;
        std::vector<double> h1(L), h2(L), gx(L), gy(L), gz(L);
        const double e = R.cfg.fd_eps;
        solid_harmonics(p, dx + e, dy, dz, false, R.cfg.flags, h1.data());
        solid_harmonics(p, dx - e, dy, dz, false, R.cfg.flags, h2.data());
        for (int i = 0; i < L; i++) gx[i] = h1[i] - h2[i];
        solid_harmonics(p, dx, dy + e, dz, false, R.cfg.flags, h1.data());
        solid_harmonics(p, dx, dy - e, dz, false, R.cfg.flags, h2.data());
        for (int i = 0; i < L; i++) gy[i] = h1[i] - h2[i];
        solid_harmonics(p, dx, dy, dz + e, false, R.cfg.flags, h1.data());
        solid_harmonics(p, dx, dy, dz - e, false, R.cfg.flags, h2.data());
        for (int i = 0; i < L; i++) gz[i] = h1[i] - h2[i];
        const double* Mx = R.M[s].data();
        const double* My = Mx + L;
        const double* Mz = My + L;
        double t[3] = {0, 0, 0};
        for (int i = 0; i < L; i++) {
            t[0] += Mz[i] * gy[i] - My[i] * gz[i];
            t[1] += Mx[i] * gz[i] - Mz[i] * gx[i];
            t[2] += My[i] * gx[i] - Mx[i] * gy[i];
        }
        for (int c = 0; c < 3; c++) out[c] += t[c] / (2 * e);
    } else if (n.first_child >= 0) {
        for (int o = 0; o < 8; o++) m2p_rot(R, S, n.first_child + o, x, out);
    }
}
// treecode M2P of the vector potential: quadrature_octree.cpp:171-194 (MAC 4R^2 < d^2, recursion
// into the children, a leaf that fails the MAC adds nothing)
static void m2p_A(const Run& R, const Tree& S, int s, const double* x, double out[3]) {
    const int p = R.cfg.p, L = nlen(p);
    const Node& n = S.nodes[s];
    double dx = x[0] - n.c[0], dy = x[1] - n.c[1], dz = x[2] - n.c[2];
    double d2 = dx * dx + dy * dy + dz * dz;
    double rad = box_radius(n);
    if (4 * rad * rad < d2) {
        if (R.M[s].empty()) return;
        std::vector<double> h(L);
        solid_harmonics(p, dx, dy, dz, false, R.cfg.flags, h.data());
        for (int c = 0; c < 3; c++) {
            double t = 0;
            for (int i = 0; i < L; i++) t += R.M[s][c * L + i] * h[i];
            out[c] += t;
        }
    } else if (n.first_child >= 0) {
        for (int o = 0; o < 8; o++) m2p_A(R, S, n.first_child + o, x, out);
    }
}
// direct vector potential of one source: q / |x - y| (the A counterpart of bs_kernel)
inline void a_kernel(const double* x, const double* y, const double q[3], double acc[3]) {
    double d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 == 0) return;
    double r = std::sqrt(r2);
    for (int c = 0; c < 3; c++) acc[c] += q[c] / r;
}
static void collect_points(const Tree& S, int s, std::vector<int>& out) {
    const Node& n = S.nodes[s];
    if (n.first_child < 0) { out.insert(out.end(), n.pts.begin(), n.pts.end()); return; }
    for (int o = 0; o < 8; o++) collect_points(S, n.first_child + o, out);
}

static Run* run_fmm(const Config& cfg, size_t ns, const double* src, const double* w, const double* J,
                    size_t nt, const double* tgt, const double* bemB = nullptr, const double* bemN = nullptr) {
    Run* R = new Run();
    R->cfg = cfg;
    const int p = cfg.p, L = nlen(p);
    R->q.resize(3 * ns);
    for (size_t i = 0; i < ns; i++) {
        double j[3];
        if (J) { j[0] = J[3 * i]; j[1] = J[3 * i + 1]; j[2] = J[3 * i + 2]; }
        else phi_hat(src + 3 * i, j);
        for (int c = 0; c < 3; c++) R->q[3 * i + c] = w[i] * j[c];
    }
    if (cfg.problem == 1) {   // BEMQuadrature: B and normal per source (bem_quadrature.hpp)
        R->bem_d.resize(ns);
        R->bem_v.resize(3 * ns);
        for (size_t i = 0; i < ns; i++) {
            const double* b = bemB + 3 * i;
            const double* n = bemN + 3 * i;
            R->bem_d[i] = w[i] * (b[0] * n[0] + b[1] * n[1] + b[2] * n[2]);
            R->bem_v[3 * i] = w[i] * (b[1] * n[2] - b[2] * n[1]);
            R->bem_v[3 * i + 1] = w[i] * (b[2] * n[0] - b[0] * n[2]);
            R->bem_v[3 * i + 2] = w[i] * (b[0] * n[1] - b[1] * n[0]);
        }
    }
    auto t0 = clk::now();
    R->S = build_tree(ns, src, cfg.source_root_c, cfg.source_root_h, cfg.cap_s, cfg.max_depth);
    double tc[3], th[3];
    if (cfg.target_root_mode == 0) bounding_box(nt, tgt, tc, th);
    else for (int a = 0; a < 3; a++) { tc[a] = cfg.target_root_c[a]; th[a] = cfg.target_root_h[a]; }
    R->T = build_tree(nt, tgt, tc, th, cfg.cap_t + 1, cfg.max_depth);
    auto t1 = clk::now();
    R->t_tree = secs(t0, t1);
    Tree& S = *R->S;
    Tree& T = *R->T;
    // P2M at useful leaves (multipole_solver.cpp:56-95)
    R->M.assign(S.nodes.size(), {});
    // OpenMP loops here and below split independent outputs only (leaves, parents of one
    // orientation, targets, levels): every sum keeps the serial order, so the results are the
    // same for any thread count
#pragma omp parallel for schedule(dynamic, 16)
    for (long long s = 0; s < (long long)S.nodes.size(); s++)
        if (is_useful(S, (int)s)) p2m_leaf(*R, S, (int)s, R->M[s]);
    auto t2 = clk::now();
    R->t_p2m = secs(t1, t2);
    // M2M, Matrices path (multipole_solver.cpp:188-229, 341-423): layers of non-empty
    // nodes; per (layer, orientation) one operator from the first node's translation.
    {
        std::vector<std::vector<int>> layers;
        for (int b : S.bfs) {
            if (!(is_sub(S, b) || is_useful(S, b))) continue;
            int lv = S.nodes[b].level;
            // enumerateNodes only reaches nodes whose ancestors are all non-empty, true here
            if ((int)layers.size() <= lv) layers.resize(lv + 1);
            layers[lv].push_back(b);
        }
        for (int b : S.bfs)
            if (is_sub(S, b) && R->M[b].empty()) R->M[b].assign(3 * L, 0.0);
        std::vector<double> Top((size_t)L * L);
        for (int lv = (int)layers.size() - 1; lv >= 1; lv--) {
            std::vector<int> by_o[8];
            for (int s : layers[lv]) {
                int par = S.nodes[s].parent;
                by_o[s - S.nodes[par].first_child].push_back(s);
            }
            for (int o = 0; o < 8; o++) {
                if (by_o[o].empty()) continue;
                int s0 = by_o[o][0], par0 = S.nodes[s0].parent;
                double t[3] = {S.nodes[s0].c[0] - S.nodes[par0].c[0], S.nodes[s0].c[1] - S.nodes[par0].c[1],
                               S.nodes[s0].c[2] - S.nodes[par0].c[2]};
                m2m_matrix_operator(p, t, cfg.flags, Top.data());
#pragma omp parallel for schedule(static)
                for (long long si = 0; si < (long long)by_o[o].size(); si++) {
                    const int s = by_o[o][si];
                    int par = S.nodes[s].parent;
                    for (int c = 0; c < 3; c++)
                        for (int i = 0; i < L; i++) {
                            double acc = 0;
                            const double* row = Top.data() + (size_t)i * L;
                            const double* x = R->M[s].data() + c * L;
                            for (int j = 0; j < L; j++) acc += row[j] * x[j];
                            R->M[par][c * L + i] += acc;
                        }
                }
            }
        }
    }
    auto t3 = clk::now();
    R->t_m2m = secs(t2, t3);
    // interaction maps (fast_multipole_solver.cpp:62-102, 105-165)
    form_maps(S, 0, T, 0, cfg.mac_scale, R->lists);
    auto t4 = clk::now();
    R->t_lists = secs(t3, t4);
    // M2L (fast_multipole_solver.cpp:258-276)
    R->Lc.assign(T.nodes.size(), {});
    for (size_t t = 0; t < T.nodes.size(); t++)
        if (is_sub(T, (int)t) || is_useful(T, (int)t)) R->Lc[t].assign(3 * L, 0.0);
    {
        // far pairs grouped by target (the set is ordered by (target, source))
        std::vector<std::pair<int, std::vector<int>>> by_t;
        for (auto& pr : R->lists.far) {
            if (by_t.empty() || by_t.back().first != pr.first) by_t.push_back({pr.first, {}});
            by_t.back().second.push_back(pr.second);
        }
#pragma omp parallel for schedule(dynamic, 4)
        for (long long g = 0; g < (long long)by_t.size(); g++) {
            std::vector<double> out(L);
            const int t = by_t[g].first;
            for (int s : by_t[g].second) {
                double tv[3] = {S.nodes[s].c[0] - T.nodes[t].c[0], S.nodes[s].c[1] - T.nodes[t].c[1],
                                S.nodes[s].c[2] - T.nodes[t].c[2]};
                Translator tr(OP_M2L, p, tv, cfg.flags);
                for (int c = 0; c < 3; c++) {
                    tr.apply(R->M[s].data() + c * L, out.data());
                    for (int i = 0; i < L; i++) R->Lc[t][c * L + i] += out[i];
                }
            }
        }
    }
    auto t5 = clk::now();
    R->t_m2l = secs(t4, t5);
    // L2L, complex path (calculation_point_octree.cpp:277-299), pre-order
    {
        // level by level (a child receives from its parent only)
        std::vector<std::vector<int>> by_level;
        for (int b : T.bfs) {
            if (!is_sub(T, b) || R->Lc[b].empty()) continue;
            int lv = T.nodes[b].level;
            if ((int)by_level.size() <= lv) by_level.resize(lv + 1);
            by_level[lv].push_back(b);
        }
        for (auto& lvl_nodes : by_level)
#pragma omp parallel for schedule(dynamic, 8)
        for (long long bi = 0; bi < (long long)lvl_nodes.size(); bi++) {
            std::vector<double> out(L);
            const int b = lvl_nodes[bi];
            for (int o = 0; o < 8; o++) {
                int ch = T.nodes[b].first_child + o;
                if (!(is_sub(T, ch) || is_useful(T, ch))) continue;
                double tv[3] = {T.nodes[b].c[0] - T.nodes[ch].c[0], T.nodes[b].c[1] - T.nodes[ch].c[1],
                                T.nodes[b].c[2] - T.nodes[ch].c[2]};
                Translator tr(OP_L2L, p, tv, cfg.flags);
                for (int c = 0; c < 3; c++) {
                    tr.apply(R->Lc[b].data() + c * L, out.data());
                    for (int i = 0; i < L; i++) R->Lc[ch][c * L + i] += out[i];
                }
            }
        }
    }
    auto t6 = clk::now();
    R->t_l2l = secs(t5, t6);
    // evaluation: calcB(current), fast_multipole_solver.cpp:214-231
    std::map<int, std::vector<int>> close_of;
    for (auto& pr : R->lists.close) close_of[pr.first].push_back(pr.second);
    const size_t nslots = T.slots.size();
    R->B.assign(3 * nslots, 0);
    R->A.assign(3 * nslots, 0);
    R->Afar.assign(3 * nslots, 0);
    R->Anear.assign(3 * nslots, 0);
    R->Bfar.assign(3 * nslots, 0);
    R->Bnear.assign(3 * nslots, 0);
    const double scale = cfg.current * MU0 / (4.0 * PI);
    std::vector<int> leaves;
    for (int b : T.bfs)
        if (is_useful(T, b)) leaves.push_back(b);
#pragma omp parallel for schedule(dynamic, 4)
    for (long long li = 0; li < (long long)leaves.size(); li++) {
        std::vector<double> h1(L), h2(L), g(3 * L);
        std::vector<std::vector<double>> gd(3, std::vector<double>(L));
        const int b = leaves[li];
        const Node& n = T.nodes[b];
        const double* Lx = R->Lc[b].data();
        const double* Ly = Lx + L;
        const double* Lz = Ly + L;
        std::vector<int> near_pts;
        if (cfg.flags & P2P) {
            auto it = close_of.find(b);
            if (it != close_of.end())
                for (int s : it->second) collect_points(S, s, near_pts);
        }
        for (int k = T.begin[T.sid2bfs[b]]; k < T.end[T.sid2bfs[b]]; k++) {
            const double* x = T.P + 3 * (size_t)T.slots[k];
            double dx = x[0] - n.c[0], dy = x[1] - n.c[1], dz = x[2] - n.c[2];
            double far[3] = {0, 0, 0};
            if (cfg.flags & ANALYTIC_CURL) {
                regular_grad_cart(p, dx, dy, dz, g.data());
                for (int i = 0; i < L; i++) {
                    far[0] += Lz[i] * g[3 * i + 1] - Ly[i] * g[3 * i + 2];
                    far[1] += Lx[i] * g[3 * i + 2] - Lz[i] * g[3 * i + 0];
                    far[2] += Ly[i] * g[3 * i + 0] - Lx[i] * g[3 * i + 1];
                }
            } else {
                const double e = cfg.fd_eps;
                double off[3][3] = {{e, 0, 0}, {0, e, 0}, {0, 0, e}};
                for (int a = 0; a < 3; a++) {
                    solid_harmonics(p, dx + off[a][0], dy + off[a][1], dz + off[a][2], true, cfg.flags, h1.data());
                    solid_harmonics(p, dx - off[a][0], dy - off[a][1], dz - off[a][2], true, cfg.flags, h2.data());
                    for (int i = 0; i < L; i++) gd[a][i] = h1[i] - h2[i];
                }
                double t[3] = {0, 0, 0};
                for (int i = 0; i < L; i++) {
                    t[0] += Lz[i] * gd[1][i] - Ly[i] * gd[2][i];
                    t[1] += Lx[i] * gd[2][i] - Lz[i] * gd[0][i];
                    t[2] += Ly[i] * gd[0][i] - Lx[i] * gd[1][i];
                }
                for (int c = 0; c < 3; c++) far[c] = t[c] / (2 * e);
            }
            double nearv[3] = {0, 0, 0};
            if (cfg.flags & P2P) {
                for (int j : near_pts) bs_kernel(x, S.P + 3 * (size_t)j, R->q.data() + 3 * (size_t)j, nearv);
            } else {
                auto it = close_of.find(b);
                if (it != close_of.end())
                    for (int s : it->second) m2p_rot(*R, S, s, x, nearv);
            }
            for (int c = 0; c < 3; c++) {
                R->Bfar[3 * k + c] = far[c] * scale;
                R->Bnear[3 * k + c] = nearv[c] * scale;
                R->B[3 * k + c] = (far[c] + nearv[c]) / (4.0 * PI) * cfg.current * MU0;
            }
            // calcA(current) (fast_multipole_solver.cpp:195-212): far part sum L R(x - c)
            // (calculation_point_octree.cpp:192-212), close nodes by M2P (quadrature_octree.cpp:
            // 171-194) or, corrected, the direct 1/|x - y| sum; scaled by current / (4 pi), no MU0
            {
                std::vector<double>& Rh = h1;
                solid_harmonics(p, dx, dy, dz, true, cfg.flags, Rh.data());
                double af[3] = {0, 0, 0}, an[3] = {0, 0, 0};
                for (int c = 0; c < 3; c++)
                    for (int i = 0; i < L; i++) af[c] += R->Lc[b][c * L + i] * Rh[i];
                if ((cfg.flags & P2P) && cfg.problem == 1) {
                    for (int j : near_pts)
                        bem_kernel(x, S.P + 3 * (size_t)j, R->bem_d[j], R->bem_v.data() + 3 * (size_t)j, an);
                } else if (cfg.flags & P2P) {
                    for (int j : near_pts) a_kernel(x, S.P + 3 * (size_t)j, R->q.data() + 3 * (size_t)j, an);
                } else {
                    auto it = close_of.find(b);
                    if (it != close_of.end())
                        for (int s : it->second) m2p_A(*R, S, s, x, an);
                }
                // calcBEM (fast_multipole_solver.cpp:233-250) divides by 4 pi only
                const double cur = cfg.problem == 1 ? 1.0 : cfg.current;
                const double sa = cur / (4.0 * PI);
                for (int c = 0; c < 3; c++) {
                    R->Afar[3 * k + c] = af[c] * sa;
                    R->Anear[3 * k + c] = an[c] * sa;
                    R->A[3 * k + c] = (af[c] + an[c]) / (4.0 * PI) * cur;
                }
            }
        }
    }
    auto t7 = clk::now();
    R->t_eval = secs(t6, t7);
    return R;
}

}  // namespace orc

// =================================================================================
// C ABI for ctypes (tests / smoke / bench cpu_baseline only)
// =================================================================================
using namespace orc;
extern "C" {

int orc_nlen(int p) { return nlen(p); }
// OpenMP threads of the parallel loops (results do not depend on the count)
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_max_threads() { return omp_get_max_threads(); }

void orc_solid_harmonics(int order, const double* v, int regular, int flags, double* out) {
    solid_harmonics(order, v[0], v[1], v[2], regular != 0, flags, out);
}
void orc_solid_harmonics_cart(int order, const double* v, int regular, double* out) {
    if (regular) regular_cart(order, v[0], v[1], v[2], out);
    else irregular_cart(order, v[0], v[1], v[2], out);
}
void orc_regular_grad(int order, const double* v, double* out) { regular_grad_cart(order, v[0], v[1], v[2], out); }

// complex M2M matrix as printed by the reference (matrices/regular_o.txt): row = output
// (l,m), column = input (l-lambda, m-mu); entries R^c(lambda,mu)(t) f(m,mu).
void orc_m2m_complex_matrix(int p, const double* t, int flags, double* re, double* im) {
    const int L = nlen(p);
    std::vector<double> Rr(L);
    std::vector<cd> R(L);
    solid_harmonics(p, t[0], t[1], t[2], true, flags, Rr.data());
    real_to_complex(p, Rr.data(), R.data());
    for (int i = 0; i < L * L; i++) { re[i] = 0; im[i] = 0; }
    for (int l = 0; l <= p; l++)
        for (int m = -l; m <= l; m++)
            for (int lambda = 0; lambda <= l; lambda++) {
                int dl = l - lambda;
                for (int mu = -lambda; mu <= lambda; mu++) {
                    int dm = m - mu;
                    if (-dl <= dm && dm <= dl) {
                        cd v = R[lm(lambda, mu)] * m2m_factor(m, mu);
                        re[(size_t)lm(l, m) * L + lm(dl, dm)] = v.real();
                        im[(size_t)lm(l, m) * L + lm(dl, dm)] = v.imag();
                    }
                }
            }
}
// real-basis translation operator T (L x L, row-major [out][in], x_out = T x_in) obtained by
// applying the reference's complex translator to each real unit vector.
void orc_operator(int kind, int p, const double* t, int flags, double* T) {
    const int L = nlen(p);
    Translator tr(kind, p, t, flags);
    std::vector<double> e(L, 0.0), out(L);
    for (int j = 0; j < L; j++) {
        e[j] = 1.0;
        tr.apply(e.data(), out.data());
        for (int i = 0; i < L; i++) T[(size_t)i * L + j] = out[i];
        e[j] = 0.0;
    }
}
void orc_m2m_matrix_operator(int p, const double* t, int flags, double* T) { m2m_matrix_operator(p, t, flags, T); }
void orc_translate(int kind, int p, const double* t, int flags, const double* in, double* out) {
    Translator tr(kind, p, t, flags);
    tr.apply(in, out);
}

// ---- trees --------------------------------------------------------------------------
void* orc_tree_build(size_t n, const double* P, int root_mode, const double* rc, const double* rh, int thr,
                     int max_depth) {
    double c[3], h[3];
    if (root_mode == 0) bounding_box(n, P, c, h);
    else for (int a = 0; a < 3; a++) { c[a] = rc[a]; h[a] = rh[a]; }
    return build_tree(n, P, c, h, thr, max_depth);
}
void orc_tree_free(void* h) { delete (Tree*)h; }
int orc_tree_num_nodes(void* h) { return (int)((Tree*)h)->nodes.size(); }
int orc_tree_num_slots(void* h) { return (int)((Tree*)h)->slots.size(); }
// arrays indexed by canonical BFS id
void orc_tree_get(void* h, int* level, int* parent, int* first_child, int* begin, int* end, int* npts,
                  double* center, double* half, int* slots) {
    Tree& T = *(Tree*)h;
    int n = (int)T.nodes.size();
    for (int b = 0; b < n; b++) {
        const Node& nd = T.nodes[T.bfs[b]];
        level[b] = nd.level;
        parent[b] = nd.parent < 0 ? -1 : T.sid2bfs[nd.parent];
        first_child[b] = nd.first_child < 0 ? -1 : T.sid2bfs[nd.first_child];
        begin[b] = T.begin[b];
        end[b] = T.end[b];
        npts[b] = (int)nd.pts.size();
        for (int a = 0; a < 3; a++) { center[3 * b + a] = nd.c[a]; half[3 * b + a] = nd.h[a]; }
    }
    for (size_t i = 0; i < T.slots.size(); i++) slots[i] = T.slots[i];
}
void orc_bounding_box(size_t n, const double* P, double* c, double* h) { bounding_box(n, P, c, h); }

// ---- interaction maps ---------------------------------------------------------------
void* orc_lists_build(void* S, void* T, double mac_scale) {
    Lists* L = new Lists();
    form_maps(*(Tree*)S, 0, *(Tree*)T, 0, mac_scale, *L);
    return L;
}
// the same with the planted tie mistake (tests of the list checker only)
void* orc_lists_build_tie_s(void* S, void* T, double mac_scale) {
    Lists* L = new Lists();
    form_maps(*(Tree*)S, 0, *(Tree*)T, 0, mac_scale, *L, true);
    return L;
}
void orc_lists_free(void* h) { delete (Lists*)h; }
void orc_lists_counts(void* h, size_t* nfar, size_t* nclose) {
    *nfar = ((Lists*)h)->far.size();
    *nclose = ((Lists*)h)->close.size();
}
// pairs as canonical BFS ids (target, source)
void orc_lists_get(void* h, void* S, void* T, int* far_t, int* far_s, int* close_t, int* close_s) {
    Lists& L = *(Lists*)h;
    Tree& s = *(Tree*)S;
    Tree& t = *(Tree*)T;
    size_t i = 0;
    for (auto& pr : L.far) { far_t[i] = t.sid2bfs[pr.first]; far_s[i] = s.sid2bfs[pr.second]; i++; }
    i = 0;
    for (auto& pr : L.close) { close_t[i] = t.sid2bfs[pr.first]; close_s[i] = s.sid2bfs[pr.second]; i++; }
}

// ---- full FMM -----------------------------------------------------------------------
struct orc_config {
    int p, cap_s, cap_t, flags, target_root_mode, max_depth;
    double target_root_c[3], target_root_h[3];
    double mac_scale, current, fd_eps;
    int problem;
    double source_root_c[3], source_root_h[3];
};
void* orc_fmm_run(const orc_config* c, size_t ns, const double* src, const double* w, const double* J, size_t nt,
                  const double* tgt) {
    Config cfg;
    cfg.p = c->p;
    cfg.cap_s = c->cap_s;
    cfg.cap_t = c->cap_t;
    cfg.flags = c->flags;
    cfg.target_root_mode = c->target_root_mode;
    cfg.max_depth = c->max_depth;
    for (int a = 0; a < 3; a++) { cfg.target_root_c[a] = c->target_root_c[a]; cfg.target_root_h[a] = c->target_root_h[a]; }
    cfg.mac_scale = c->mac_scale;
    cfg.current = c->current;
    cfg.fd_eps = c->fd_eps > 0 ? c->fd_eps : FD_EPS;
    cfg.problem = c->problem;
    for (int a = 0; a < 3; a++) { cfg.source_root_c[a] = c->source_root_c[a]; cfg.source_root_h[a] = c->source_root_h[a]; }
    return run_fmm(cfg, ns, src, w, J, nt, tgt);
}
// Problem::BEM: sources with B and normal vectors (BEMQuadrature)
void* orc_fmm_run_bem(const orc_config* c, size_t ns, const double* src, const double* w, const double* Bv,
                      const double* nrm, size_t nt, const double* tgt) {
    Config cfg;
    cfg.p = c->p;
    cfg.cap_s = c->cap_s;
    cfg.cap_t = c->cap_t;
    cfg.flags = c->flags;
    cfg.target_root_mode = c->target_root_mode;
    cfg.max_depth = c->max_depth;
    for (int a = 0; a < 3; a++) { cfg.target_root_c[a] = c->target_root_c[a]; cfg.target_root_h[a] = c->target_root_h[a]; }
    cfg.mac_scale = c->mac_scale;
    cfg.current = 1.0;
    cfg.fd_eps = c->fd_eps > 0 ? c->fd_eps : FD_EPS;
    cfg.problem = 1;
    for (int a = 0; a < 3; a++) { cfg.source_root_c[a] = c->source_root_c[a]; cfg.source_root_h[a] = c->source_root_h[a]; }
    return run_fmm(cfg, ns, src, w, nullptr, nt, tgt, Bv, nrm);
}
// math::calcBEMIntegral (integration.cpp:204-215): sum w BEMFunction / (4 pi)
void orc_direct_bem(size_t ns, const double* src, const double* w, const double* Bv, const double* nrm, size_t nt,
                    const double* tgt, double* out) {
    std::vector<double> d(ns), v(3 * ns);
    for (size_t i = 0; i < ns; i++) {
        const double* b = Bv + 3 * i;
        const double* n = nrm + 3 * i;
        d[i] = w[i] * (b[0] * n[0] + b[1] * n[1] + b[2] * n[2]);
        v[3 * i] = w[i] * (b[1] * n[2] - b[2] * n[1]);
        v[3 * i + 1] = w[i] * (b[2] * n[0] - b[0] * n[2]);
        v[3 * i + 2] = w[i] * (b[0] * n[1] - b[1] * n[0]);
    }
#pragma omp parallel for schedule(static)
    for (long long t = 0; t < (long long)nt; t++) {
        double acc[3] = {0, 0, 0};
        for (size_t i = 0; i < ns; i++) bem_kernel(tgt + 3 * t, src + 3 * i, d[i], v.data() + 3 * i, acc);
        for (int c = 0; c < 3; c++) out[3 * t + c] = acc[c] / (4 * PI);
    }
}
void orc_fmm_free(void* h) { delete (Run*)h; }
int orc_fmm_num_slots(void* h) { return (int)((Run*)h)->T->slots.size(); }
void orc_fmm_get(void* h, int* slots, double* B, double* Bfar, double* Bnear) {
    Run& R = *(Run*)h;
    size_t n = R.T->slots.size();
    for (size_t i = 0; i < n; i++) slots[i] = R.T->slots[i];
    std::memcpy(B, R.B.data(), 3 * n * sizeof(double));
    std::memcpy(Bfar, R.Bfar.data(), 3 * n * sizeof(double));
    std::memcpy(Bnear, R.Bnear.data(), 3 * n * sizeof(double));
}
void orc_fmm_get_A(void* h, double* A, double* Afar, double* Anear) {
    Run& R = *(Run*)h;
    size_t n = R.T->slots.size();
    std::memcpy(A, R.A.data(), 3 * n * sizeof(double));
    std::memcpy(Afar, R.Afar.data(), 3 * n * sizeof(double));
    std::memcpy(Anear, R.Anear.data(), 3 * n * sizeof(double));
}
// treecode single-point evaluation over the run's source tree from node bfs_id (0 = root):
// MultipoleSolver::calcB(current, point) = root calcRot * current MU0 / (4 pi) and
// calcA(current, point) = root calcA * current / (4 pi) (multipole_solver.cpp:686-700)
void orc_treecode(void* h, int bfs_id, size_t n, const double* x, double* B, double* A) {
    Run& R = *(Run*)h;
    const int s = R.S->bfs[bfs_id];
#pragma omp parallel for schedule(dynamic, 16)
    for (long long i = 0; i < (long long)n; i++) {
        double b[3] = {0, 0, 0}, a[3] = {0, 0, 0};
        m2p_rot(R, *R.S, s, x + 3 * i, b);
        m2p_A(R, *R.S, s, x + 3 * i, a);
        for (int c = 0; c < 3; c++) {
            B[3 * i + c] = b[c] / (4.0 * PI) * R.cfg.current * MU0;
            A[3 * i + c] = a[c] / (4.0 * PI) * R.cfg.current;
        }
    }
}
void orc_fmm_times(void* h, double* t) {
    Run& R = *(Run*)h;
    t[0] = R.t_tree; t[1] = R.t_p2m; t[2] = R.t_m2m; t[3] = R.t_lists; t[4] = R.t_m2l; t[5] = R.t_l2l; t[6] = R.t_eval;
}
void orc_fmm_counts(void* h, size_t* c) {
    Run& R = *(Run*)h;
    c[0] = R.S->nodes.size(); c[1] = R.T->nodes.size(); c[2] = R.lists.far.size(); c[3] = R.lists.close.size();
}
// local expansion (3*L, component-major) of a target node given by canonical BFS id
void orc_fmm_local(void* h, int bfs_id, double* out) {
    Run& R = *(Run*)h;
    int s = R.T->bfs[bfs_id];
    const int L = nlen(R.cfg.p);
    if (R.Lc[s].empty()) { for (int i = 0; i < 3 * L; i++) out[i] = 0; return; }
    std::memcpy(out, R.Lc[s].data(), 3 * L * sizeof(double));
}
void orc_fmm_multipole(void* h, int bfs_id, double* out) {
    Run& R = *(Run*)h;
    int s = R.S->bfs[bfs_id];
    const int L = nlen(R.cfg.p);
    if (R.M[s].empty()) { for (int i = 0; i < 3 * L; i++) out[i] = 0; return; }
    std::memcpy(out, R.M[s].data(), 3 * L * sizeof(double));
}

// ---- direct sums --------------------------------------------------------------------
// B(x) = MU0/(4 PI) current sum_q w_q J_q x (x - y_q)/|x - y_q|^3  (integration.cpp:101-120).
// J == NULL means J = phi_hat(y).
void orc_direct_B(size_t ns, const double* src, const double* w, const double* J, size_t nt, const double* tgt,
                  double current, double* B) {
    std::vector<double> q(3 * ns);
    for (size_t i = 0; i < ns; i++) {
        double j[3];
        if (J) { j[0] = J[3 * i]; j[1] = J[3 * i + 1]; j[2] = J[3 * i + 2]; }
        else phi_hat(src + 3 * i, j);
        for (int c = 0; c < 3; c++) q[3 * i + c] = w[i] * j[c];
    }
    const double scale = MU0 / (4 * PI) * current;
#pragma omp parallel for schedule(static)
    for (long long t = 0; t < (long long)nt; t++) {
        double acc[3] = {0, 0, 0};
        for (size_t i = 0; i < ns; i++) bs_kernel(tgt + 3 * t, src + 3 * i, q.data() + 3 * i, acc);
        for (int c = 0; c < 3; c++) B[3 * t + c] = acc[c] * scale;
    }
}
// A(x) = current / (4 PI) sum_q w_q J_q / |x - y_q| (the calcA scaling of the reference:
// multipole_solver.cpp:686-692, fast_multipole_solver.cpp:195-212; no MU0)
void orc_direct_A(size_t ns, const double* src, const double* w, const double* J, size_t nt, const double* tgt,
                  double current, double* A) {
    std::vector<double> q(3 * ns);
    for (size_t i = 0; i < ns; i++) {
        double j[3];
        if (J) { j[0] = J[3 * i]; j[1] = J[3 * i + 1]; j[2] = J[3 * i + 2]; }
        else phi_hat(src + 3 * i, j);
        for (int c = 0; c < 3; c++) q[3 * i + c] = w[i] * j[c];
    }
    const double scale = current / (4 * PI);
#pragma omp parallel for schedule(static)
    for (long long t = 0; t < (long long)nt; t++) {
        double acc[3] = {0, 0, 0};
        for (size_t i = 0; i < ns; i++) a_kernel(tgt + 3 * t, src + 3 * i, q.data() + 3 * i, acc);
        for (int c = 0; c < 3; c++) A[3 * t + c] = acc[c] * scale;
    }
}
// dL/dw_j for L = sum_i |B_i - Bstar_i|^2 with B from the direct sum.
void orc_direct_grad_w(size_t ns, const double* src, const double* w, const double* J, size_t nt, const double* tgt,
                       double current, const double* Bstar, double* grad) {
    std::vector<double> B(3 * nt);
    orc_direct_B(ns, src, w, J, nt, tgt, current, B.data());
    const double scale = MU0 / (4 * PI) * current;
#pragma omp parallel for schedule(static)
    for (long long j = 0; j < (long long)ns; j++) {
        double Jj[3];
        if (J) { Jj[0] = J[3 * j]; Jj[1] = J[3 * j + 1]; Jj[2] = J[3 * j + 2]; }
        else phi_hat(src + 3 * j, Jj);
        double acc = 0;
        for (size_t i = 0; i < nt; i++) {
            double kv[3] = {0, 0, 0};
            bs_kernel(tgt + 3 * i, src + 3 * j, Jj, kv);
            for (int c = 0; c < 3; c++) acc += 2 * (B[3 * i + c] - Bstar[3 * i + c]) * kv[c] * scale;
        }
        grad[j] = acc;
    }
}

}  // extern "C"

// Particle <-> expansion kernels: source charges, P2M, L2P (curl of the local
// expansion from analytic gradients), near-field P2P, final assembly, and the FP64
// direct Biot-Savart sum (math::calcBioSavartLaplace, integration.cpp:101-120).
//
// Harmonics use the Cartesian recurrences of the complex regular solid harmonics
// Phi_l^m = r^l P_l^m(cos) e^{im phi}/(l+m)! (m >= 0); the reference's real basis is
// R(l,0) = (-1)^l Phi_l^0, R(l,m) = (-1)^l sqrt2 Re Phi_l^m, R(l,-m) = (-1)^l sqrt2 Im Phi_l^m
// (harmonics.cpp:20-56 with the sectoral term m = p included).  All recurrences run in
// FP64 on the scaled vector u = (x - c) / sigma; only P2P runs in FP32 (with FP64
// accumulation per target).
#include <algorithm>

#include "fmm_expand.hpp"
#include "fmm_recip.hpp"

namespace fmm {

__device__ __forceinline__ int lmi(int l, int m) { return l * l + l + m; }
__device__ __forceinline__ double sgnl(int l) { return (l & 1) ? -1.0 : 1.0; }
constexpr double SQRT2 = 1.41421356237309504880;

// ---------------------------------------------------------------- charges
__global__ void k_charges(const int* __restrict__ slots, int n, const double* __restrict__ xyz,
                          const double* __restrict__ w, const double* __restrict__ J, double* __restrict__ sxyz,
                          double* __restrict__ sq, double* __restrict__ sJ) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = slots[k];
    double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    double j0, j1, j2;
    if (J) {
        j0 = J[3 * i]; j1 = J[3 * i + 1]; j2 = J[3 * i + 2];
    } else {
        // current direction perp().normalized() (vector3.cu:154-167, 201-204)
        double px = -y, py = x;
        double l = sqrt(px * px + py * py);
        j0 = l != 0 ? px / l : 0.0;
        j1 = l != 0 ? py / l : 0.0;
        j2 = 0.0;
    }
    double ww = w[i];
    sxyz[3 * k] = x; sxyz[3 * k + 1] = y; sxyz[3 * k + 2] = z;
    sq[3 * k] = ww * j0; sq[3 * k + 1] = ww * j1; sq[3 * k + 2] = ww * j2;
    if (sJ) { sJ[3 * k] = j0; sJ[3 * k + 1] = j1; sJ[3 * k + 2] = j2; }
}

__global__ void k_gather_xyz(const int* __restrict__ slots, int n, const double* __restrict__ xyz,
                             double* __restrict__ out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = slots[k];
    out[3 * k] = xyz[3 * i]; out[3 * k + 1] = xyz[3 * i + 1]; out[3 * k + 2] = xyz[3 * i + 2];
}

// ---------------------------------------------------------------- P2M
// M_l^{+-m} = sum_j q_j R_l^{+-m}(u_j), u = (y - c) / sigma, with R(l, 0) = (-1)^l Phi_l^0 and
// R(l, +-m) = (-1)^l sqrt2 {Re, Im} Phi_l^m (harmonics.cpp:281-320).  One wave (a one-wave
// workgroup) per useful source leaf.  The leaf's tasks (point, m) -- point-major, P + 1 per point
// -- run 64 at a time; task (point, m) makes column m of the harmonics in one loop over the
// degree l = 1 .. P (P, the order, is a template parameter, so the loop unrolls with l fixed):
// for l <= m the sectoral step Phi_l^l = Phi_{l-1}^{l-1} (x + iy) / (2l), for l > m the degree
// step Phi_l^m = ((2l - 1) z Phi_{l-1}^m - r^2 Phi_{l-2}^m) / (l^2 - m^2), both as
// (A Phi_{l-1} - b Phi_{l-2}) f with a complex A = (x + iy or (2l - 1) z), b = (0 up to l = m + 1,
// then r^2) and f from an LDS table F[m][l] (1/(2l) for l <= m, 1/(l^2 - m^2) above), so every
// lane runs the same P steps whatever its m; from l = m on each value is written to the point's
// row of a small LDS ring.  After each round the points it completed are added, lane k owning
// coefficients k, k + 64, ... (ring rows padded to whole slots of 64, so the reads need no
// masks), in point order with FP64 fused multiply-adds; the degree signs and sqrt2 are applied
// once when the leaf is written.  The ring holds the points of one round and its two boundary
// points (64 / (P + 1) + 2 rows).  T is the type of the recurrences and the ring: FP64 in the
// FP64 mode, FP32 otherwise (|u| <= sqrt3 keeps every Phi_l^m in range and the recurrences
// forward-stable).
template <class T> __device__ __forceinline__ T recip_t(int n);
template <> __device__ __forceinline__ double recip_t<double>(int n) { return recip(n); }
template <> __device__ __forceinline__ float recip_t<float>(int n) { return recipf(n); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <int P>
constexpr int p2m_ring() { return 64 / (P + 1) + 2; }
template <int P>
constexpr int p2m_row() { return 64 * (((P + 1) * (P + 1) + 63) / 64); }

template <class T, int P>
__global__ __launch_bounds__(64) void k_p2m(const int* __restrict__ leaves, const double* __restrict__ center,
                                            const int* __restrict__ begin, const int* __restrict__ end,
                                            const int* __restrict__ level, ScaleTable sig,
                                            const double* __restrict__ sxyz, const double* __restrict__ sq, int Lp,
                                            T* __restrict__ M, const float* __restrict__ dm, int Kp,
                                            float* __restrict__ cscale, _Float16* __restrict__ hl) {
    constexpr int P1 = P + 1, L = P1 * P1, NS = (L + 63) / 64, RS = p2m_row<P>(), RING = p2m_ring<P>();
    __shared__ T ring[RING * RS];   // complex Phi of the points in flight (Re at (l, m), Im at (l, -m))
    __shared__ T F[P1 * P1];        // F[m][l]: the step factor of column m at degree l
    const int lane = threadIdx.x;
    for (int i = lane; i < P1 * P1; i += 64) {
        const int m = i / P1, l = i - m * P1;
        F[i] = l == 0 ? (T)0 : l <= m ? recip_t<T>(2 * l) : recip_t<T>(l * l - m * m);
    }
    const int leaf = leaves[blockIdx.x];
    const double cx = center[3 * leaf], cy = center[3 * leaf + 1], cz = center[3 * leaf + 2];
    const double inv = 1.0 / sig.s[level[leaf]];
    const int jb = begin[leaf], n = end[leaf] - jb, ntask = n * P1;
    double acc[NS][3];
#pragma unroll
    for (int r = 0; r < NS; r++) acc[r][0] = acc[r][1] = acc[r][2] = 0.0;
    // this lane's task: point pt, column m; a round moves it 64 tasks on
    int pt = lane / P1, m = lane - pt * P1;
    constexpr int dpt = 64 / P1, dmv = 64 - dpt * P1;
    int done = 0;
    __syncthreads();   // F is written
    for (int t0 = 0; t0 < ntask; t0 += 64) {
        if (t0 + lane < ntask) {
            const int j = jb + pt;
            const T ux = (T)((sxyz[3 * j] - cx) * inv), uy = (T)((sxyz[3 * j + 1] - cy) * inv),
                    uz = (T)((sxyz[3 * j + 2] - cz) * inv);
            const T r2 = ux * ux + uy * uy + uz * uz;
            T* row = ring + (pt % RING) * RS;
            const T* Fm = F + m * P1;
            T p1r = 1, p1i = 0, p2r = 0, p2i = 0;
            if (m == 0) row[0] = (T)1;
#pragma unroll
            for (int l = 1; l <= P; l++) {
                const bool sect = l <= m;
                const T ar = sect ? ux : (T)(2 * l - 1) * uz, ai = sect ? uy : (T)0;
                const T b = l <= m + 1 ? (T)0 : r2, f = Fm[l];
                const T vr = fma_t(ar, p1r, fma_t(-ai, p1i, -(b * p2r))) * f;
                const T vi = fma_t(ar, p1i, fma_t(ai, p1r, -(b * p2i))) * f;
                p2r = p1r; p2i = p1i; p1r = vr; p1i = vi;
                if (l >= m) {
                    row[l * l + l + m] = vr;
                    if (m > 0) row[l * l + l - m] = vi;
                }
            }
        }
        pt += dpt;
        m += dmv;
        if (m >= P1) { m -= P1; pt++; }
        __syncthreads();   // one wave: the ring rows are written before they are read
        const int complete = min(n, (t0 + 64) / P1);
        for (int q = done; q < complete; q++) {
            const T* row = ring + (q % RING) * RS + lane;
            const double q0 = sq[3 * (jb + q)], q1 = sq[3 * (jb + q) + 1], q2 = sq[3 * (jb + q) + 2];
            T v[NS];
#pragma unroll
            for (int r = 0; r < NS; r++) v[r] = row[64 * r];   // beyond L: padding, never written out
#pragma unroll
            for (int r = 0; r < NS; r++) {
                const double x = (double)v[r];
                acc[r][0] = fma_t(q0, x, acc[r][0]);
                acc[r][1] = fma_t(q1, x, acc[r][1]);
                acc[r][2] = fma_t(q2, x, acc[r][2]);
            }
        }
        done = complete;
        __syncthreads();   // the rows are read before the next round overwrites them
    }
    T* out = M + (size_t)leaf * 3 * Lp;
    T mv[NS][3];   // the leaf's multipoles as written (zero beyond L)
#pragma unroll
    for (int r = 0; r < NS; r++) {
        const int k = lane + 64 * r;
        if (k < L) {
            const int l = (int)sqrtf((float)k + 0.5f), mm = k - l * l - l;   // k = l^2 + l + mm
            const double s = ((l & 1) ? -1.0 : 1.0) * (mm == 0 ? 1.0 : SQRT2);
            for (int cc = 0; cc < 3; cc++) mv[r][cc] = (T)(acc[r][cc] * s);
        } else {
            for (int cc = 0; cc < 3; cc++) mv[r][cc] = (T)0;
        }
        if (k < Lp)
            for (int cc = 0; cc < 3; cc++) out[cc * Lp + k] = mv[r][cc];
    }
    for (int k = lane + 64 * NS; k < Lp; k += 64)   // padding beyond what the loop above covers
        for (int cc = 0; cc < 3; cc++) out[cc * Lp + k] = (T)0;
    if constexpr (sizeof(T) == 4) {
        // F16X3: the leaf's three columns of the M2L's B operand as k_split_columns makes them
        // (the same FP32 operations: dm-scaled coefficients, power-of-2 column scale, FP16 hi and
        // lo, [column][K step][hi | lo]), so that the split after the M2M covers the other nodes
        if (hl) {
            float dmv[NS];
#pragma unroll
            for (int r = 0; r < NS; r++) {
                const int k = lane + 64 * r;
                dmv[r] = k < L && k < Kp ? dm[k] : 0.0f;
            }
            for (int cc = 0; cc < 3; cc++) {
                float x[NS], mx = 0.0f;
#pragma unroll
                for (int r = 0; r < NS; r++) {
                    x[r] = mv[r][cc] * dmv[r];
                    mx = fmaxf(mx, fabsf(x[r]));
                }
                for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
                const float sc = mx > 0.0f ? exp2f((float)(-(ilogbf(mx) + 1))) : 1.0f;
                const size_t col = (size_t)leaf * 3 + cc;
                if (lane == 0) cscale[col] = sc;
                _Float16* dst = hl + col * 2 * Kp;
#pragma unroll
                for (int r = 0; r < NS; r++) {
                    const int k = lane + 64 * r;
                    if (k < Kp) {
                        const float y = x[r] * sc;
                        const _Float16 h = (_Float16)y;
                        dst[(k >> 5) * 64 + (k & 31)] = h;
                        dst[(k >> 5) * 64 + (k & 31) + 32] = (_Float16)(y - (float)h);
                    }
                }
            }
        }
    }
}

// The nodes k_p2m does not write (subdivided ones and leaves without points), cleared with
// 16-byte stores, one wave per node (n4 vectors per node).
__global__ __launch_bounds__(256) void k_zero_unwritten(const int* __restrict__ first_child,
                                                        const int* __restrict__ npts, int nnodes, int n4,
                                                        float4* __restrict__ M) {
    const int v = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (v >= nnodes || (first_child[v] < 0 && npts[v] > 0)) return;
    float4* o = M + (size_t)v * n4;
    for (int i = lane; i < n4; i += 64) o[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---------------------------------------------------------------- L2P
// One wave per useful target leaf; lane (m, g) evaluates column m for the targets g, g+G, ...
// (G = 64 / p) and the p lanes of a target are summed through LDS.  The local expansion is turned into the complex
// coefficients beta_c (degree <= p-1) of B = sigma^-2 Re sum beta_c(l,m) Phi_l^m(u), using
// dz Phi_l^m = Phi_{l-1}^m, (dx - i dy) Phi_l^m = Phi_{l-1}^{m-1}, (dx + i dy) Phi_l^m = -Phi_{l-1}^{m+1}.
__device__ __forceinline__ void alpha_of(const double* Ls, int L, int c, int l, int m, double& ar, double& ai) {
    const double* Lc = Ls + c * L;
    double s = sgnl(l);
    if (m == 0) { ar = s * Lc[lmi(l, 0)]; ai = 0.0; }
    else { ar = s * SQRT2 * Lc[lmi(l, m)]; ai = -s * SQRT2 * Lc[lmi(l, -m)]; }
}
// derivative coefficients of component c along axis a (0 x, 1 y, 2 z) at (l', m')
__device__ __forceinline__ void deriv_coef(const double* Ls, int L, int c, int a, int lp, int mp, double& br,
                                           double& bi) {
    int l = lp + 1;
    double r, i;
    if (a == 2) { alpha_of(Ls, L, c, l, mp, br, bi); return; }
    double sr = 0, si = 0;
    alpha_of(Ls, L, c, l, mp + 1, r, i);
    double ur = 0.5 * r, ui = 0.5 * i;            // 1/2 alpha(l, m'+1)
    double vr = 0, vi = 0;                        // 1/2 alpha(l, m'-1) (+ the m' = 1 extra)
    if (mp >= 1) {
        alpha_of(Ls, L, c, l, mp - 1, r, i);
        vr = 0.5 * r; vi = 0.5 * i;
        if (mp == 1) { vr *= 2.0; vi *= 2.0; }
    }
    if (a == 0) { sr = ur - vr; si = ui - vi; }   // x: 1/2 a(m'+1) - 1/2 a(m'-1)
    else {                                        // y: (i/2) (a(m'+1) + a(m'-1))
        double wr = ur + vr, wi = ui + vi;
        sr = -wi; si = wr;
    }
    br = sr; bi = si;
}

// POT: the vector potential A = sigma^-1 Re sum alpha_c(l,m) Phi_l^m(u) over degree <= p (the
// reference's calcA, calculation_point_octree.cpp:192-212) instead of B = curl A (degree <= p-1).
//
// One wave (a one-wave workgroup) per useful target leaf.  The leaf's local expansion is mapped
// once to the coefficients beta (LDS, type T); then each lane evaluates one target: the sectoral
// chain Phi_m^m = Phi_{m-1}^{m-1} (x + iy) / (2m) and, per m, Phi_l^m = ((2l - 1) z Phi_{l-1}^m -
// r^2 Phi_{l-2}^m) / (l^2 - m^2), every (l, m) term added to the three components as it is made.
// D (the highest degree evaluated) is a template parameter, so the loops unroll with the
// recurrence factors as constants and the term index and the beta offsets fixed; the beta reads
// are the same address in every lane (LDS broadcast); explicit fused multiply-adds.  T is FP64
// in the FP64 mode, FP32 otherwise (|u| <= sqrt3 inside the leaf box keeps the recurrences in
// range); the result is scaled and written in FP64.
__device__ __forceinline__ float fma_l(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_l(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <class T, int D, bool POT>
__global__ __launch_bounds__(64) void k_l2p(const int* __restrict__ leaves, const double* __restrict__ center,
                                            const int* __restrict__ begin, const int* __restrict__ end,
                                            const int* __restrict__ level, ScaleTable sig,
                                            const double* __restrict__ txyz, const void* __restrict__ Lexp_v, int p,
                                            int Lp, bool lexp_f64, double* __restrict__ Bfar) {
    constexpr int NB = (D + 1) * (D + 2) / 2;   // (l', m') pairs with l' <= D
    extern __shared__ __attribute__((aligned(16))) double sh[];
    const int L = (p + 1) * (p + 1);
    double* Ls = sh;                             // [3][L] the local expansion (FP64)
    T* beta = (T*)(sh + 3 * L);                  // [NB][3][2]
    const int leaf = leaves[blockIdx.x];
    const int lane = threadIdx.x;
    if (lexp_f64) {
        const double* Lsrc = (const double*)Lexp_v + (size_t)leaf * 3 * Lp;
        for (int e = lane; e < 3 * L; e += 64) {
            const int c = e / L, k = e - c * L;
            Ls[e] = Lsrc[c * Lp + k];
        }
    } else {
        const float* Lsrc = (const float*)Lexp_v + (size_t)leaf * 3 * Lp;
        for (int e = lane; e < 3 * L; e += 64) {
            const int c = e / L, k = e - c * L;
            Ls[e] = (double)Lsrc[c * Lp + k];
        }
    }
    __syncthreads();
    for (int e = lane; e < NB; e += 64) {
        int lp = 0;
        while ((lp + 1) * (lp + 2) / 2 <= e) lp++;
        const int mp = e - lp * (lp + 1) / 2;
        double bv[3][2];
        if (POT) {
            for (int c = 0; c < 3; c++) alpha_of(Ls, L, c, lp, mp, bv[c][0], bv[c][1]);
        } else {
            double d[3][3][2];   // d[comp][axis]
            for (int c = 0; c < 3; c++)
                for (int a = 0; a < 3; a++) deriv_coef(Ls, L, c, a, lp, mp, d[c][a][0], d[c][a][1]);
            for (int q = 0; q < 2; q++) {   // B = curl A
                bv[0][q] = d[2][1][q] - d[1][2][q];
                bv[1][q] = d[0][2][q] - d[2][0][q];
                bv[2][q] = d[1][0][q] - d[0][1][q];
            }
        }
        for (int c = 0; c < 3; c++) {
            beta[(e * 3 + c) * 2] = (T)bv[c][0];
            beta[(e * 3 + c) * 2 + 1] = (T)bv[c][1];
        }
    }
    __syncthreads();
    const double cx = center[3 * leaf], cy = center[3 * leaf + 1], cz = center[3 * leaf + 2];
    const double inv = 1.0 / sig.s[level[leaf]];
    const double s2 = POT ? inv : inv * inv;
    const int kb = begin[leaf], n = end[leaf] - kb;
    for (int t0 = 0; t0 < n; t0 += 64) {
        const int t = t0 + lane;
        if (t >= n) break;
        const int k = kb + t;
        const T ux = (T)((txyz[3 * k] - cx) * inv), uy = (T)((txyz[3 * k + 1] - cy) * inv),
                uz = (T)((txyz[3 * k + 2] - cz) * inv);
        const T r2 = ux * ux + uy * uy + uz * uz;
        T b0 = 0, b1 = 0, b2 = 0;
        T sr = 1, si = 0;   // Phi_m^m
#pragma unroll
        for (int m = 0; m <= D; m++) {
            if (m > 0) {
                constexpr T one = 1;
                const T h = one / (T)(2 * m);
                const T nr = (sr * ux - si * uy) * h, ni = (sr * uy + si * ux) * h;
                sr = nr;
                si = ni;
            }
            T p1r = sr, p1i = si, p2r = 0, p2i = 0;
#pragma unroll
            for (int l = m; l <= D; l++) {
                T vr = p1r, vi = p1i;
                if (l > m) {
                    const T a = (T)(2 * l - 1) * uz, f = (T)1 / (T)(l * l - m * m);
                    vr = fma_l(a, p1r, -r2 * p2r) * f;
                    vi = fma_l(a, p1i, -r2 * p2i) * f;
                    p2r = p1r; p2i = p1i; p1r = vr; p1i = vi;
                }
                const T* bt = beta + (l * (l + 1) / 2 + m) * 6;
                b0 = fma_l(bt[0], vr, fma_l(-bt[1], vi, b0));
                b1 = fma_l(bt[2], vr, fma_l(-bt[3], vi, b1));
                b2 = fma_l(bt[4], vr, fma_l(-bt[5], vi, b2));
            }
        }
        Bfar[3 * k] = (double)b0 * s2;
        Bfar[3 * k + 1] = (double)b1 * s2;
        Bfar[3 * k + 2] = (double)b2 * s2;
    }
}

// ---------------------------------------------------------------- P2P
// Near field B(x) = sum_j q_j x (x - y_j) / |x - y_j|^3 over the sources of each close
// source node (the close-list loop of FastMultipoleSolver::calcB, fast_multipole_solver.cpp:214-229;
// a close node may be internal, then all its subtree's sources are summed).
//
// Positions are taken relative to one origin o (the source root centre).  For R = float both
// sides are unevaluated sums hi + lo of two floats formed in FP64 (~2^-48 of the root size),
// so the pair difference d = (x_hi - y_hi) + (x_lo - y_lo) stays accurate for very close
// pairs (x_hi - y_hi is exact when the two are within a factor 2) while the pair arithmetic
// runs in FP32.  R = double keeps the offsets whole.
template <class R> struct V4;
template <> struct V4<float> {
    typedef float4 t;
    static __device__ float4 make(float a, float b, float c, float d) { return make_float4(a, b, c, d); }
};
template <> struct V4<double> {
    typedef double4 t;
    static __device__ double4 make(double a, double b, double c, double d) { return make_double4(a, b, c, d); }
};
// v_rsq_f32 directly (1 ulp): pair distances squared are far above the denormal range, so the
// denormal-scaling wrapper of rsqrtf is not needed
__device__ __forceinline__ float rsqrt_t(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsqrt_t(double x) { return 1.0 / sqrt(x); }
// P0[j] = (hi(y_j - o), q_x), P1[j] = (lo(y_j - o), q_y), P2[j] = q_z for every source slot j
template <class R>
__global__ void k_pack_sources(int n, const double* __restrict__ origin, const double* __restrict__ sxyz,
                               const double* __restrict__ sq, typename V4<R>::t* __restrict__ P0,
                               typename V4<R>::t* __restrict__ P1, R* __restrict__ P2) {
    constexpr bool split = sizeof(R) == 4;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double dx = sxyz[3 * j] - origin[0], dy = sxyz[3 * j + 1] - origin[1], dz = sxyz[3 * j + 2] - origin[2];
    R hx = (R)dx, hy = (R)dy, hz = (R)dz;
    R lx = split ? (R)(dx - (double)hx) : (R)0, ly = split ? (R)(dy - (double)hy) : (R)0;
    R lz = split ? (R)(dz - (double)hz) : (R)0;
    P0[j] = V4<R>::make(hx, hy, hz, (R)sq[3 * j]);
    P1[j] = V4<R>::make(lx, ly, lz, (R)sq[3 * j + 1]);
    P2[j] = (R)sq[3 * j + 2];
}

// One workgroup (4 waves) per target leaf.  The 64-source chunks of the leaf's close list are
// dealt round-robin to the waves; each wave stages a chunk in its own LDS slice and its lanes
// sweep it, G = floor(64 / n) lanes per target for a leaf of 8 <= n <= 64 targets (64 / 2^ceil(log2 n)
// below 8; lane = target * G + g takes sources g, g + G, ...).  Lane partials are summed over
// the G lanes, then over the 4 waves in a fixed order, and written once.
// POT: the vector potential sum_j q_j / |x - y_j| instead of B.
template <class R, bool POT, int UNROLL>
__device__ __forceinline__ void p2p_leaf(int leaf, const int* __restrict__ tbegin, const int* __restrict__ tend,
                                         const int* __restrict__ close_begin, const int* __restrict__ close_end,
                                         const int* __restrict__ close_s, const double* __restrict__ origin,
                                         const int* __restrict__ sbegin, const int* __restrict__ send,
                                         const typename V4<R>::t* __restrict__ P0,
                                         const typename V4<R>::t* __restrict__ P1, const R* __restrict__ P2,
                                         const double* __restrict__ txyz, double* __restrict__ Bnear,
                                         typename V4<R>::t (*s0)[64], typename V4<R>::t (*s1)[64], R (*s2)[64],
                                         double (*red)[3][64]) {
    typedef typename V4<R>::t v4;
    constexpr bool split = sizeof(R) == 4;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tb = tbegin[leaf], te = tend[leaf];
    const int cb = close_begin[leaf], ce = close_end[leaf];
    if (cb >= ce) {
        for (int k = tb + tid; k < te; k += 256) Bnear[3 * k] = Bnear[3 * k + 1] = Bnear[3 * k + 2] = 0.0;
        return;
    }
    for (int k0 = tb; k0 < te; k0 += 64) {
        const int n = min(64, te - k0);
        // G lanes per target: 64 / n (n >= 8, at most 8 lanes, summed through LDS), else the power
        // of 2 64 / 2^ceil(log2 n) (summed by shuffles)
        const bool pow2 = n < 8;
        const int lgG = 6 - (n <= 1 ? 0 : 32 - __clz(n - 1));
        const int G = pow2 ? 1 << lgG : 64 / n;
        const int ti = pow2 ? lane >> lgG : lane / G, g = lane - ti * G;
        const bool valid = ti < n;
        double ux = 0, uy = 0, uz = 0;
        if (valid) {
            ux = txyz[3 * (k0 + ti)] - origin[0];
            uy = txyz[3 * (k0 + ti) + 1] - origin[1];
            uz = txyz[3 * (k0 + ti) + 2] - origin[2];
        }
        const R uhx = (R)ux, uhy = (R)uy, uhz = (R)uz;
        const R ulx = split ? (R)(ux - (double)uhx) : (R)0, uly = split ? (R)(uy - (double)uhy) : (R)0;
        const R ulz = split ? (R)(uz - (double)uhz) : (R)0;
        double a0 = 0, a1 = 0, a2 = 0;
        // the close list's 64-source chunks, in list order, dealt round-robin to the waves (an
        // internal close node can hold most of a leaf's sources)
        int run = 0;   // chunks of the close nodes before ci
        for (int ci = cb; ci < ce; ci++) {
            const int s = close_s[ci];
            const int sb = sbegin[s], se = send[s];
            const int nch = (se - sb + 63) >> 6;
            for (int j0 = sb + 64 * ((wave - run) & 3); j0 < se; j0 += 4 * 64) {
                const int nj = min(64, se - j0);
                __syncwarp();
                if (lane < nj) {
                    s0[wave][lane] = P0[j0 + lane];
                    s1[wave][lane] = P1[j0 + lane];
                    s2[wave][lane] = P2[j0 + lane];
                }
                __syncwarp();
                R f0 = 0, f1 = 0, f2 = 0;
#pragma unroll UNROLL
                for (int jj = g; jj < nj; jj += G) {
                    const v4 yh = s0[wave][jj];
                    const v4 yl = s1[wave][jj];
                    const R qx = yh.w, qy = yl.w, qz = s2[wave][jj];
                    R dx = uhx - yh.x, dy = uhy - yh.y, dz = uhz - yh.z;
                    if (split) {
                        dx += ulx - yl.x;
                        dy += uly - yl.y;
                        dz += ulz - yl.z;
                    }
                    // fused multiply-adds (explicit: this library is built without contraction)
                    const R r2 = fma_t(dx, dx, fma_t(dy, dy, dz * dz));
                    const R ri = r2 > (R)0 ? rsqrt_t(r2) : (R)0;
                    if (POT) {
                        f0 = fma_t(qx, ri, f0);
                        f1 = fma_t(qy, ri, f1);
                        f2 = fma_t(qz, ri, f2);
                    } else {
                        const R ri3 = ri * ri * ri;
                        f0 = fma_t(fma_t(qy, dz, -qz * dy), ri3, f0);
                        f1 = fma_t(fma_t(qz, dx, -qx * dz), ri3, f1);
                        f2 = fma_t(fma_t(qx, dy, -qy * dx), ri3, f2);
                    }
                }
                a0 += f0;
                a1 += f1;
                a2 += f2;
            }
            run += nch;
        }
        if (pow2)
            for (int off = 1; off < G; off <<= 1) {
                a0 += __shfl_xor(a0, off);
                a1 += __shfl_xor(a1, off);
                a2 += __shfl_xor(a2, off);
            }
        red[wave][0][lane] = a0;
        red[wave][1][lane] = a1;
        red[wave][2][lane] = a2;
        __syncthreads();
        if (wave == 0 && valid && g == 0) {
            const int k = k0 + ti;
            for (int c = 0; c < 3; c++) {
                double t[4];
                for (int w = 0; w < 4; w++) {
                    double sw = red[w][c][lane];
                    if (!pow2)
                        for (int gg = 1; gg < G; gg++) sw += red[w][c][lane + gg];
                    t[w] = sw;
                }
                Bnear[3 * k + c] = ((t[0] + t[1]) + t[2]) + t[3];
            }
        }
        __syncthreads();
    }
}

template <class R, bool POT>
__global__ __launch_bounds__(256) void k_p2p(const int* __restrict__ leaves, const int* __restrict__ tbegin,
                                             const int* __restrict__ tend, const int* __restrict__ close_begin,
                                             const int* __restrict__ close_end, const int* __restrict__ close_s,
                                             const double* __restrict__ origin, const int* __restrict__ sbegin,
                                             const int* __restrict__ send, const typename V4<R>::t* __restrict__ P0,
                                             const typename V4<R>::t* __restrict__ P1, const R* __restrict__ P2,
                                             const double* __restrict__ txyz, double* __restrict__ Bnear) {
    __shared__ typename V4<R>::t s0[4][64];
    __shared__ typename V4<R>::t s1[4][64];
    __shared__ R s2[4][64];
    __shared__ double red[4][3][64];
    p2p_leaf<R, POT, 4>(leaves[blockIdx.x], tbegin, tend, close_begin, close_end, close_s, origin, sbegin, send, P0, P1, P2, txyz, Bnear, s0, s1, s2, red);
}

// The same near field as a small persistent grid taking leaves from a queue (an atomic counter
// that starts at 0), with a register budget of 56 per lane: one workgroup per CU fits beside the
// F16X3 M2L GEMM's, so the near field runs on the VALU slots the GEMM leaves idle.
template <class R, bool POT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(10))) void k_p2p_queue(
    const int* __restrict__ leaves, int nleaves, int* __restrict__ queue, const int* __restrict__ tbegin,
    const int* __restrict__ tend, const int* __restrict__ close_begin, const int* __restrict__ close_end,
    const int* __restrict__ close_s, const double* __restrict__ origin, const int* __restrict__ sbegin,
    const int* __restrict__ send, const typename V4<R>::t* __restrict__ P0, const typename V4<R>::t* __restrict__ P1,
    const R* __restrict__ P2, const double* __restrict__ txyz, double* __restrict__ Bnear) {
    __shared__ typename V4<R>::t s0[4][64];
    __shared__ typename V4<R>::t s1[4][64];
    __shared__ R s2[4][64];
    __shared__ double red[4][3][64];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(queue, 1);
        __syncthreads();
        const int i = next;
        __syncthreads();   // every thread has read next before thread 0 takes the following leaf
        if (i >= nleaves) break;
        p2p_leaf<R, POT, 1>(leaves[i], tbegin, tend, close_begin, close_end, close_s, origin, sbegin, send, P0, P1, P2, txyz, Bnear, s0, s1, s2, red);
    }
}

// ---------------------------------------------------------------- Problem::BEM
// BEMQuadrature sources (bem_quadrature.hpp): weight w, field B and normal n per point.  The
// source strengths in slot order are d = w (B . n) and v = w (B x n) (3); their multipoles are
// M_c(l, m) = -sum_q (dR_lm/dy_c d_q + (grad R_lm x v_q)_c) (integration.cpp:295-353, with the
// exact gradient in place of the reference's central differences), the potential they give is
// sum_q ((y - x) d + (y - x) x v) / |y - x|^3 (integration.cpp:191-202).
__global__ void k_bem_charges(const int* __restrict__ slots, int n, const double* __restrict__ xyz,
                              const double* __restrict__ w, const double* __restrict__ Bv,
                              const double* __restrict__ nrm, double* __restrict__ sxyz, double* __restrict__ sv,
                              double* __restrict__ sd) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = slots[k];
    const double b0 = Bv[3 * i], b1 = Bv[3 * i + 1], b2 = Bv[3 * i + 2];
    const double n0 = nrm[3 * i], n1 = nrm[3 * i + 1], n2 = nrm[3 * i + 2], ww = w[i];
    sxyz[3 * k] = xyz[3 * i]; sxyz[3 * k + 1] = xyz[3 * i + 1]; sxyz[3 * k + 2] = xyz[3 * i + 2];
    sd[k] = ww * (b0 * n0 + b1 * n1 + b2 * n2);
    sv[3 * k] = ww * (b1 * n2 - b2 * n1);
    sv[3 * k + 1] = ww * (b2 * n0 - b0 * n2);
    sv[3 * k + 2] = ww * (b0 * n1 - b1 * n0);
}

// One workgroup (64 threads) per useful source leaf.  Points go through LDS in chunks of
// BEM_CHUNK: thread t builds the complex table Phi_n^k (n <= p - 1, k >= 0) of point t; then each
// thread owns coefficients i = tid + 64 r and adds the gradient of R_i at every point of the chunk:
// dz Phi_l^m = Phi_{l-1}^m, dx Phi_l^m = (Phi_{l-1}^{m-1} - Phi_{l-1}^{m+1}) / 2,
// dy Phi_l^m = i (Phi_{l-1}^{m-1} + Phi_{l-1}^{m+1}) / 2, Phi_n^{-k} = (-1)^k conj Phi_n^k.
constexpr int BEM_CHUNK = 8;
template <class T>
__global__ __launch_bounds__(64) void k_p2m_bem(const int* __restrict__ leaves, const double* __restrict__ center,
                                                const int* __restrict__ begin, const int* __restrict__ end,
                                                const int* __restrict__ level, ScaleTable sig,
                                                const double* __restrict__ sxyz, const double* __restrict__ sv,
                                                const double* __restrict__ sd, int p, int Lp, T* __restrict__ M) {
    extern __shared__ double2 tab[];   // [BEM_CHUNK][p (p + 1) / 2]
    __shared__ double str[BEM_CHUNK][4];
    const int NT = p * (p + 1) / 2, L = (p + 1) * (p + 1);
    const int leaf = leaves[blockIdx.x], tid = threadIdx.x;
    const double cx = center[3 * leaf], cy = center[3 * leaf + 1], cz = center[3 * leaf + 2];
    const double inv = 1.0 / sig.s[level[leaf]];
    constexpr int RMAX = (MAX_ORDER + 1) * (MAX_ORDER + 1) / 64 + 1;
    double acc[RMAX][3];
    for (int r = 0; r < RMAX; r++) acc[r][0] = acc[r][1] = acc[r][2] = 0.0;
    const int jb = begin[leaf], je = end[leaf];
    for (int j0 = jb; j0 < je; j0 += BEM_CHUNK) {
        const int nj = min(BEM_CHUNK, je - j0);
        __syncthreads();
        if (tid < nj) {
            const int j = j0 + tid;
            const double ux = (sxyz[3 * j] - cx) * inv, uy = (sxyz[3 * j + 1] - cy) * inv,
                         uz = (sxyz[3 * j + 2] - cz) * inv, r2 = ux * ux + uy * uy + uz * uz;
            double2* t = tab + tid * NT;
            double dr = 1.0, di = 0.0;   // Phi_k^k = (x + iy)^k / (2^k k!)
            for (int k = 0; k < p; k++) {
                if (k > 0) {
                    const double s = recip(2 * k);
                    const double nr = (dr * ux - di * uy) * s, ni = (dr * uy + di * ux) * s;
                    dr = nr; di = ni;
                }
                double p1r = dr, p1i = di, p2r = 0.0, p2i = 0.0;
                t[k * (k + 1) / 2 + k] = make_double2(dr, di);
                for (int n = k + 1; n < p; n++) {
                    const double f = recip(n * n - k * k);
                    const double vr = ((2.0 * n - 1) * uz * p1r - r2 * p2r) * f;
                    const double vi = ((2.0 * n - 1) * uz * p1i - r2 * p2i) * f;
                    p2r = p1r; p2i = p1i; p1r = vr; p1i = vi;
                    t[n * (n + 1) / 2 + k] = make_double2(vr, vi);
                }
            }
            str[tid][0] = sd[j] * inv;
            str[tid][1] = sv[3 * j] * inv;
            str[tid][2] = sv[3 * j + 1] * inv;
            str[tid][3] = sv[3 * j + 2] * inv;
        }
        __syncthreads();
        for (int r = 0; r < RMAX; r++) {
            const int i = tid + 64 * r;
            if (i >= L) break;
            int l = 0;
            while ((l + 1) * (l + 1) <= i) l++;
            const int m = i - l * l - l, am = m < 0 ? -m : m;
            if (l == 0) continue;
            const int n = l - 1;
            const double sg = ((l & 1) ? -1.0 : 1.0) * (am == 0 ? 1.0 : 1.41421356237309504880);
            for (int q = 0; q < nj; q++) {
                const double2* t = tab + q * NT;
                // Phi_{l-1}^{k} for k = am - 1, am, am + 1 (signed orders through the conjugate rule)
                auto ph = [&](int k) -> double2 {
                    if (k > n || k < -n) return make_double2(0.0, 0.0);
                    if (k >= 0) return t[n * (n + 1) / 2 + k];
                    const double2 v = t[n * (n + 1) / 2 - k];
                    const double s = (k & 1) ? -1.0 : 1.0;
                    return make_double2(s * v.x, -s * v.y);
                };
                const double2 lo = ph(am - 1), mid = ph(am), hi = ph(am + 1);
                const double gxr = 0.5 * (lo.x - hi.x), gxi = 0.5 * (lo.y - hi.y);
                const double gyr = -0.5 * (lo.y + hi.y), gyi = 0.5 * (lo.x + hi.x);
                const double gzr = mid.x, gzi = mid.y;
                const bool im = m < 0;
                const double gx = sg * (im ? gxi : gxr), gy = sg * (im ? gyi : gyr), gz = sg * (im ? gzi : gzr);
                const double d = str[q][0], v0 = str[q][1], v1 = str[q][2], v2 = str[q][3];
                acc[r][0] -= gx * d + (gy * v2 - gz * v1);
                acc[r][1] -= gy * d + (gz * v0 - gx * v2);
                acc[r][2] -= gz * d + (gx * v1 - gy * v0);
            }
        }
    }
    T* out = M + (size_t)leaf * 3 * Lp;
    for (int e = tid; e < 3 * Lp; e += 64) out[e] = (T)0;
    __syncthreads();
    for (int r = 0; r < RMAX; r++) {
        const int i = tid + 64 * r;
        if (i >= L) break;
        for (int c = 0; c < 3; c++) out[c * Lp + i] = (T)acc[r][c];
    }
}

// near field of the BEM potential, FP64: one thread per target of a useful target leaf, over the
// sources of every node of the leaf's close list
__global__ __launch_bounds__(64) void k_p2p_bem(const int* __restrict__ leaves, const int* __restrict__ tbegin,
                                                const int* __restrict__ tend, const int* __restrict__ close_begin,
                                                const int* __restrict__ close_end, const int* __restrict__ close_s,
                                                const int* __restrict__ sbegin, const int* __restrict__ send,
                                                const double* __restrict__ sxyz, const double* __restrict__ sv,
                                                const double* __restrict__ sd, const double* __restrict__ txyz,
                                                double* __restrict__ out) {
    const int leaf = leaves[blockIdx.x];
    for (int k = tbegin[leaf] + threadIdx.x; k < tend[leaf]; k += blockDim.x) {
        const double x = txyz[3 * k], y = txyz[3 * k + 1], z = txyz[3 * k + 2];
        double a0 = 0, a1 = 0, a2 = 0;
        for (int ci = close_begin[leaf]; ci < close_end[leaf]; ci++) {
            const int s = close_s[ci];
            for (int j = sbegin[s]; j < send[s]; j++) {
                const double q0 = sxyz[3 * j] - x, q1 = sxyz[3 * j + 1] - y, q2 = sxyz[3 * j + 2] - z;
                const double r2 = q0 * q0 + q1 * q1 + q2 * q2;
                if (r2 == 0.0) continue;
                const double r3 = 1.0 / (r2 * sqrt(r2)), d = sd[j];
                const double v0 = sv[3 * j], v1 = sv[3 * j + 1], v2 = sv[3 * j + 2];
                a0 += (q0 * d + (q1 * v2 - q2 * v1)) * r3;
                a1 += (q1 * d + (q2 * v0 - q0 * v2)) * r3;
                a2 += (q2 * d + (q0 * v1 - q1 * v0)) * r3;
            }
        }
        out[3 * k] = a0;
        out[3 * k + 1] = a1;
        out[3 * k + 2] = a2;
    }
}

void launch_bem_charges(const int* slots, int n, const double* xyz, const double* w, const double* Bv,
                        const double* nrm, double* sxyz, double* sv, double* sd, hipStream_t st) {
    if (n) k_bem_charges<<<blocks(n), 256, 0, st>>>(slots, n, xyz, w, Bv, nrm, sxyz, sv, sd);
    FMM_HIP(hipGetLastError());
}
void launch_p2m_bem(bool f64, const int* leaves, int nleaves, const TreeDev& S, const ScaleTable& sig,
                    const double* sxyz, const double* sv, const double* sd, int p, int Lp, void* M, hipStream_t st) {
    if (!nleaves) return;
    if (p < 1 || p > MAX_ORDER) throw Error(1, "BEM P2M: order out of range");
    const size_t lds = (size_t)BEM_CHUNK * (p * (p + 1) / 2) * sizeof(double2);
    if (f64)
        k_p2m_bem<double><<<nleaves, 64, lds, st>>>(leaves, S.center.as<double>(), S.begin.as<int>(), S.end.as<int>(),
                                                    S.level.as<int>(), sig, sxyz, sv, sd, p, Lp, (double*)M);
    else
        k_p2m_bem<float><<<nleaves, 64, lds, st>>>(leaves, S.center.as<double>(), S.begin.as<int>(), S.end.as<int>(),
                                                   S.level.as<int>(), sig, sxyz, sv, sd, p, Lp, (float*)M);
    FMM_HIP(hipGetLastError());
}
void launch_p2p_bem(const int* leaves, int nleaves, const TreeDev& T, const TreeDev& S, const int* close_begin,
                    const int* close_end, const int* close_s, const double* sxyz, const double* sv, const double* sd,
                    const double* txyz, double* out, hipStream_t st) {
    if (!nleaves) return;
    k_p2p_bem<<<nleaves, 64, 0, st>>>(leaves, T.begin.as<int>(), T.end.as<int>(), close_begin, close_end, close_s,
                                      S.begin.as<int>(), S.end.as<int>(), sxyz, sv, sd, txyz, out);
    FMM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- assembly
__global__ void k_assemble(const int* __restrict__ slots, int n, const double* __restrict__ Bfar,
                           const double* __restrict__ Bnear, double scale, double* __restrict__ B) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = slots[k];
    for (int c = 0; c < 3; c++) B[3 * i + c] = scale * (Bfar[3 * k + c] + Bnear[3 * k + c]);
}

// ---------------------------------------------------------------- direct sum (FP64)
// B(x) = scale sum_j q_j x (x - y_j) / |x - y_j|^3 (POT: A(x) = scale sum_j q_j / |x - y_j|),
// math::calcBioSavartLaplace (integration.cpp:101-120).  The sources are cut into nsplit
// contiguous ranges: workgroup (b, s) sums range s for the 256 targets of block b (sources staged
// in LDS) into partial[s][t][3]; k_direct_reduce adds the nsplit partials of each target in range
// order (deterministic) and scales.  Splitting the sources keeps every CU busy when the targets
// are few (16,384 targets are 64 workgroups).
template <bool POT>
__global__ __launch_bounds__(256) void k_direct_partial(const double* __restrict__ sxyz, const double* __restrict__ sq,
                                                        int ns, const double* __restrict__ txyz, int nt,
                                                        double* __restrict__ partial) {
    __shared__ double3 sp[256];
    __shared__ double3 sqq[256];
    const int t = blockIdx.x * blockDim.x + threadIdx.x, split = blockIdx.y, nsplit = gridDim.y;
    const long long per = ((long long)ns + nsplit - 1) / nsplit;
    const int j_begin = (int)min((long long)ns, per * split), j_end = (int)min((long long)ns, per * (split + 1));
    double x = 0, y = 0, z = 0;
    if (t < nt) { x = txyz[3 * t]; y = txyz[3 * t + 1]; z = txyz[3 * t + 2]; }
    double a0 = 0, a1 = 0, a2 = 0;
    for (int j0 = j_begin; j0 < j_end; j0 += 256) {
        int j = j0 + threadIdx.x;
        __syncthreads();
        if (j < j_end) {
            sp[threadIdx.x] = make_double3(sxyz[3 * j], sxyz[3 * j + 1], sxyz[3 * j + 2]);
            sqq[threadIdx.x] = make_double3(sq[3 * j], sq[3 * j + 1], sq[3 * j + 2]);
        }
        __syncthreads();
        int nj = min(256, j_end - j0);
        for (int jj = 0; jj < nj; jj++) {
            double dx = x - sp[jj].x, dy = y - sp[jj].y, dz = z - sp[jj].z;
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 == 0.0) continue;
            double r = sqrt(r2);
            double3 q = sqq[jj];
            if (POT) {
                double ri = 1.0 / r;
                a0 += q.x * ri;
                a1 += q.y * ri;
                a2 += q.z * ri;
            } else {
                double r3 = 1.0 / (r2 * r);
                a0 += (q.y * dz - q.z * dy) * r3;
                a1 += (q.z * dx - q.x * dz) * r3;
                a2 += (q.x * dy - q.y * dx) * r3;
            }
        }
    }
    if (t < nt) {
        double* o = partial + ((size_t)split * nt + t) * 3;
        o[0] = a0;
        o[1] = a1;
        o[2] = a2;
    }
}

__global__ void k_direct_reduce(const double* __restrict__ partial, int nt, int nsplit, double scale,
                                double* __restrict__ B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;   // target * 3 + component
    if (i >= 3 * nt) return;
    double acc = 0;
    for (int s = 0; s < nsplit; s++) acc += partial[(size_t)s * 3 * nt + i];
    B[i] = acc * scale;
}

int direct_splits(int ns, int nt) {
    // about 4 workgroups per CU in total, each range at least 4,096 sources
    const int tb = std::max(1, blocks(nt));
    int s = std::max(1, 4 * num_cus() / tb);
    s = std::min(s, std::max(1, ns / 4096));
    return s;
}

void launch_direct(bool pot, const double* sxyz, const double* sq, int ns, const double* txyz, int nt, double scale,
                   double* out, double* partial, int nsplit, hipStream_t st) {
    if (!nt) return;
    if (nsplit < 1) throw Error(1, "direct sum: at least one source range");
    dim3 grid(blocks(nt), nsplit);
    if (pot) k_direct_partial<true><<<grid, 256, 0, st>>>(sxyz, sq, ns, txyz, nt, partial);
    else k_direct_partial<false><<<grid, 256, 0, st>>>(sxyz, sq, ns, txyz, nt, partial);
    k_direct_reduce<<<blocks(3 * (long long)nt), 256, 0, st>>>(partial, nt, nsplit, scale, out);
    FMM_HIP(hipGetLastError());
}

void launch_charges(const int* slots, int n, const double* xyz, const double* w, const double* J, double* sxyz,
                    double* sq, double* sJ, hipStream_t st) {
    if (n) k_charges<<<blocks(n), 256, 0, st>>>(slots, n, xyz, w, J, sxyz, sq, sJ);
}
void launch_gather_xyz(const int* slots, int n, const double* xyz, double* out, hipStream_t st) {
    if (n) k_gather_xyz<<<blocks(n), 256, 0, st>>>(slots, n, xyz, out);
}
void launch_zero_unwritten(const TreeDev& S, int Lp, size_t esize, void* M, hipStream_t st) {
    const size_t bytes = (size_t)3 * Lp * esize;
    if (!S.num_nodes) return;
    if (bytes % 16) throw Error(1, "expansion rows must be whole 16-byte vectors");
    k_zero_unwritten<<<(S.num_nodes + 3) / 4, 256, 0, st>>>(S.first_child.as<int>(), S.npts.as<int>(), S.num_nodes,
                                                            (int)(bytes / 16), (float4*)M);
    FMM_HIP(hipGetLastError());
}
template <int P>
static void p2m_go(bool f64, const int* leaves, int nleaves, const TreeDev& S, const ScaleTable& sig,
                   const double* sxyz, const double* sq, int Lp, void* M, const F16Leaves& split, hipStream_t st) {
    if (f64)
        k_p2m<double, P><<<nleaves, 64, 0, st>>>(leaves, S.center.as<double>(), S.begin.as<int>(), S.end.as<int>(),
                                                 S.level.as<int>(), sig, sxyz, sq, Lp, (double*)M, nullptr, 0,
                                                 nullptr, nullptr);
    else
        k_p2m<float, P><<<nleaves, 64, 0, st>>>(leaves, S.center.as<double>(), S.begin.as<int>(), S.end.as<int>(),
                                                S.level.as<int>(), sig, sxyz, sq, Lp, (float*)M, split.dm, split.Kp,
                                                split.cscale, split.hl);
}
void launch_p2m(bool f64, const int* leaves, int nleaves, const TreeDev& S, const ScaleTable& sig, const double* sxyz,
                const double* sq, int p, int Lp, void* M, hipStream_t st, const F16Leaves& split) {
    if (!nleaves) return;
    if (p < 0 || p > MAX_ORDER) throw Error(4, "expansion order above 20");
    if (split.hl && (f64 || split.Kp % 32 || split.Kp > Lp || !split.dm || !split.cscale))
        throw Error(4, "P2M: FP16 split of the leaves needs FP32 multipoles and 32 | Kp <= Lp");
    switch (p) {
#define FMM_P2M_P(PV) \
    case PV: p2m_go<PV>(f64, leaves, nleaves, S, sig, sxyz, sq, Lp, M, split, st); break;
        FMM_P2M_P(0) FMM_P2M_P(1) FMM_P2M_P(2) FMM_P2M_P(3) FMM_P2M_P(4) FMM_P2M_P(5) FMM_P2M_P(6) FMM_P2M_P(7)
        FMM_P2M_P(8) FMM_P2M_P(9) FMM_P2M_P(10) FMM_P2M_P(11) FMM_P2M_P(12) FMM_P2M_P(13) FMM_P2M_P(14)
        FMM_P2M_P(15) FMM_P2M_P(16) FMM_P2M_P(17) FMM_P2M_P(18) FMM_P2M_P(19) FMM_P2M_P(20)
#undef FMM_P2M_P
        default: throw Error(4, "expansion order above 20");
    }
    FMM_HIP(hipGetLastError());
}
template <class T, int D, bool POT>
static void l2p_go(const int* leaves, int nleaves, const TreeDev& Tr, const ScaleTable& sig, const double* txyz,
                   const void* Lexp, bool lexp_f64, int p, int Lp, double* out, hipStream_t st) {
    constexpr int NB = (D + 1) * (D + 2) / 2;
    const size_t lds = (size_t)3 * (p + 1) * (p + 1) * sizeof(double) + (size_t)NB * 6 * sizeof(T);
    k_l2p<T, D, POT><<<nleaves, 64, lds, st>>>(leaves, Tr.center.as<double>(), Tr.begin.as<int>(), Tr.end.as<int>(),
                                               Tr.level.as<int>(), sig, txyz, Lexp, p, Lp, lexp_f64, out);
}
template <int D, bool POT>
static void l2p_d(bool f64, const int* leaves, int nleaves, const TreeDev& Tr, const ScaleTable& sig,
                  const double* txyz, const void* Lexp, int p, int Lp, double* out, hipStream_t st) {
    if (f64)
        l2p_go<double, D, POT>(leaves, nleaves, Tr, sig, txyz, Lexp, true, p, Lp, out, st);
    else
        l2p_go<float, D, POT>(leaves, nleaves, Tr, sig, txyz, Lexp, false, p, Lp, out, st);
}
template <bool POT>
static void l2p_launch(bool f64, const int* leaves, int nleaves, const TreeDev& T, const ScaleTable& sig,
                       const double* txyz, const void* Lexp, int p, int Lp, double* out, hipStream_t st) {
    const int D = POT ? p : p - 1;
    if (p < 1 || p > MAX_ORDER) throw Error(1, "L2P: order out of range");
    if (!nleaves) return;
    switch (D) {
#define FMM_L2P_D(DV) \
    case DV: l2p_d<DV, POT>(f64, leaves, nleaves, T, sig, txyz, Lexp, p, Lp, out, st); break;
        FMM_L2P_D(0) FMM_L2P_D(1) FMM_L2P_D(2) FMM_L2P_D(3) FMM_L2P_D(4) FMM_L2P_D(5) FMM_L2P_D(6) FMM_L2P_D(7)
        FMM_L2P_D(8) FMM_L2P_D(9) FMM_L2P_D(10) FMM_L2P_D(11) FMM_L2P_D(12) FMM_L2P_D(13) FMM_L2P_D(14)
        FMM_L2P_D(15) FMM_L2P_D(16) FMM_L2P_D(17) FMM_L2P_D(18) FMM_L2P_D(19) FMM_L2P_D(20)
#undef FMM_L2P_D
        default: throw Error(1, "L2P: order out of range");
    }
    FMM_HIP(hipGetLastError());
}
void launch_l2p(bool f64, const int* leaves, int nleaves, const TreeDev& T, const ScaleTable& sig, const double* txyz,
                const void* Lexp, int p, int Lp, double* Bfar, hipStream_t st) {
    if (!nleaves) return;
    l2p_launch<false>(f64, leaves, nleaves, T, sig, txyz, Lexp, p, Lp, Bfar, st);
}
void launch_l2p_A(bool f64, const int* leaves, int nleaves, const TreeDev& T, const ScaleTable& sig, const double* txyz,
                  const void* Lexp, int p, int Lp, double* Afar, hipStream_t st) {
    if (!nleaves) return;
    l2p_launch<true>(f64, leaves, nleaves, T, sig, txyz, Lexp, p, Lp, Afar, st);
}
void launch_pack_sources(bool f64, int n, const double* origin, const double* sxyz, const double* sq, void* P0,
                         void* P1, void* P2, hipStream_t st) {
    if (!n) return;
    if (f64)
        k_pack_sources<double><<<blocks(n), 256, 0, st>>>(n, origin, sxyz, sq, (double4*)P0, (double4*)P1, (double*)P2);
    else
        k_pack_sources<float><<<blocks(n), 256, 0, st>>>(n, origin, sxyz, sq, (float4*)P0, (float4*)P1, (float*)P2);
    FMM_HIP(hipGetLastError());
}
template <bool POT>
static void p2p_launch(bool f64, const int* leaves, int nleaves, const TreeDev& T, const TreeDev& S,
                       const int* close_begin, const int* close_end, const int* close_s, const void* P0,
                       const void* P1, const void* P2, const double* txyz, double* out, hipStream_t st) {
    if (!nleaves) return;
    if (f64)
        k_p2p<double, POT><<<nleaves, 256, 0, st>>>(leaves, T.begin.as<int>(), T.end.as<int>(),
                                                    close_begin, close_end, close_s, S.center.as<double>(),
                                                    S.begin.as<int>(), S.end.as<int>(), (const double4*)P0,
                                                    (const double4*)P1, (const double*)P2, txyz, out);
    else
        k_p2p<float, POT><<<nleaves, 256, 0, st>>>(leaves, T.begin.as<int>(), T.end.as<int>(),
                                                   close_begin, close_end, close_s, S.center.as<double>(),
                                                   S.begin.as<int>(), S.end.as<int>(), (const float4*)P0,
                                                   (const float4*)P1, (const float*)P2, txyz, out);
    FMM_HIP(hipGetLastError());
}
void launch_p2p(bool f64, const int* leaves, int nleaves, const TreeDev& T, const TreeDev& S, const int* close_begin,
                const int* close_end, const int* close_s, const void* P0, const void* P1, const void* P2,
                const double* txyz, double* Bnear, hipStream_t st) {
    p2p_launch<false>(f64, leaves, nleaves, T, S, close_begin, close_end, close_s, P0, P1, P2, txyz, Bnear, st);
}
void launch_p2p_queue(bool f64, const int* leaves, int nleaves, const TreeDev& T, const TreeDev& S,
                      const int* close_begin, const int* close_end, const int* close_s, const void* P0, const void* P1,
                      const void* P2, const double* txyz, double* Bnear, int* queue, int grid, hipStream_t st) {
    if (!nleaves) return;
    if (grid < 1) throw Error(1, "near field: empty persistent grid");
    FMM_HIP(hipMemsetAsync(queue, 0, sizeof(int), st));
    if (f64)
        k_p2p_queue<double, false><<<grid, 256, 0, st>>>(leaves, nleaves, queue, T.begin.as<int>(), T.end.as<int>(),
                                                         close_begin, close_end, close_s, S.center.as<double>(),
                                                         S.begin.as<int>(), S.end.as<int>(), (const double4*)P0,
                                                         (const double4*)P1, (const double*)P2, txyz, Bnear);
    else
        k_p2p_queue<float, false><<<grid, 256, 0, st>>>(leaves, nleaves, queue, T.begin.as<int>(), T.end.as<int>(),
                                                        close_begin, close_end, close_s, S.center.as<double>(),
                                                        S.begin.as<int>(), S.end.as<int>(), (const float4*)P0,
                                                        (const float4*)P1, (const float*)P2, txyz, Bnear);
    FMM_HIP(hipGetLastError());
}
void launch_p2p_A(bool f64, const int* leaves, int nleaves, const TreeDev& T, const TreeDev& S, const int* close_begin,
                  const int* close_end, const int* close_s, const void* P0, const void* P1, const void* P2,
                  const double* txyz, double* Anear, hipStream_t st) {
    p2p_launch<true>(f64, leaves, nleaves, T, S, close_begin, close_end, close_s, P0, P1, P2, txyz, Anear, st);
}
void launch_assemble(const int* slots, int n, const double* Bfar, const double* Bnear, double scale, double* B,
                     hipStream_t st) {
    if (n) k_assemble<<<blocks(n), 256, 0, st>>>(slots, n, Bfar, Bnear, scale, B);
}

}  // namespace fmm

// Adjoint (transposed) FMM: gradient of L(w) = sum_i |B_i(w) - B*_i|^2 with respect to the
// per-source current weights, given g_i = dL/dB_i.  B is linear in the source strengths
// q_j = w_j J_j, so dL/dq_j is the transposed FMM applied to g:
//   L2P^T : per target leaf, P = sum_i g_i R(u_i) (a P2M of the targets with charges g,
//           order p-1), then the transposed curl map gives lambda (a local-expansion shaped
//           adjoint);
//   L2L^T : bottom-up, lambda_parent += T_L2L^T lambda_child;
//   M2L^T : mu_s += T_M2L^T lambda_t over all far pairs;
//   M2M^T : top-down, mu_child += T_M2M^T mu_parent;
//   P2M^T : dL/dq_j = sum_lm mu(l,m) R_lm(u_j);
//   P2P^T : dL/dq_j += sum_{close targets i} (x_i - y_j) x g_i / |x_i - y_j|^3;
// and dL/dw_j = J_j . dL/dq_j.  The backward pass consumes no forward expansion (B is
// linear in w), so nothing per level is stored for it: the same expansion-sized
// workspaces are reused, and only the trees, lists and operator tables of the forward
// pass are read.  The transposed GEMMs reuse the forward tiles with source and
// destination swapped and the operator read transposed.
#include "fmm_adjoint.hpp"
#include "fmm_expand.hpp"

namespace fmm {

__device__ __forceinline__ int lmi(int l, int m) { return l * l + l + m; }
__device__ __forceinline__ double sgnl(int l) { return (l & 1) ? -1.0 : 1.0; }
constexpr double SQRT2_A = 1.41421356237309504880;
constexpr double RSQRT2_A = 0.70710678118654752440;

__global__ void k_gather_g(const int* __restrict__ slots, int n, const double* __restrict__ dLdB, double scale,
                           double* __restrict__ g) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = slots[k];
    for (int c = 0; c < 3; c++) g[3 * k + c] = scale * dLdB[3 * i + c];
}

// gamma_b(l', m') from the real-basis P_b of degree l' <= p-1
__device__ __forceinline__ void gamma_of(const double* P, int Lp, int b, int lp, int mp, double& gr, double& gi) {
    const double* Pb = P + b * Lp;
    if (lp < 0 || mp > lp || mp < 0) { gr = 0; gi = 0; return; }
    double s = sgnl(lp);
    if (mp == 0) { gr = s * Pb[lmi(lp, 0)]; gi = 0; }
    else { gr = s * RSQRT2_A * Pb[lmi(lp, mp)]; gi = s * RSQRT2_A * Pb[lmi(lp, -mp)]; }
}

// Gamma_c(l, m) contribution of axis a with curl output b and sign coef (see header)
__device__ __forceinline__ void axis_terms(const double* P, int Lp, int p, int a, int b, double coef, int l, int m,
                                           double& Gr, double& Gi) {
    int lp = l - 1;
    if (lp > p - 1) return;
    double gr, gi;
    if (a == 2) {   // z: Gamma(l, m) += coef gamma_b(l-1, m)
        if (m <= lp) { gamma_of(P, Lp, b, lp, m, gr, gi); Gr += coef * gr; Gi += coef * gi; }
        return;
    }
    // terms from e = (l-1, m-1) [alpha(l, m'+1) with m' = m-1] and e = (l-1, m+1) [alpha(l, m'-1) with m' = m+1]
    double ar = 0, ai = 0, br = 0, bi = 0;
    if (m >= 1) gamma_of(P, Lp, b, lp, m - 1, ar, ai);
    if (m + 1 <= lp) gamma_of(P, Lp, b, lp, m + 1, br, bi);
    double f = (m == 0) ? 2.0 : 1.0;
    if (a == 0) {   // x: +1/2 gamma(m-1) - f/2 gamma(m+1)
        Gr += coef * (0.5 * ar - 0.5 * f * br);
        Gi += coef * (0.5 * ai - 0.5 * f * bi);
    } else {        // y: (i/2) gamma(m-1) + (i f/2) gamma(m+1)
        double sr = 0.5 * ar + 0.5 * f * br, si = 0.5 * ai + 0.5 * f * bi;
        Gr += coef * (-si);
        Gi += coef * sr;
    }
}

// lambda (adjoint local expansion) of each useful target leaf from its P expansion
template <class T>
__global__ void k_l2p_adj(const int* __restrict__ leaves, const int* __restrict__ level, ScaleTable sig,
                          const double* __restrict__ Pexp, int p, int Lp, T* __restrict__ lam) {
    const int leaf = leaves[blockIdx.x];
    const double* P = Pexp + (size_t)leaf * 3 * Lp;
    T* out = lam + (size_t)leaf * 3 * Lp;
    const double s2 = 1.0 / (sig.s[level[leaf]] * sig.s[level[leaf]]);
    const int NT = (p + 1) * (p + 2) / 2;
    for (int e = threadIdx.x; e < NT; e += blockDim.x) {
        int l = 0;
        while ((l + 1) * (l + 2) / 2 <= e) l++;
        int m = e - l * (l + 1) / 2;
        double G[3][2] = {{0, 0}, {0, 0}, {0, 0}};
        // B_x = dy A_z - dz A_y ; B_y = dz A_x - dx A_z ; B_z = dx A_y - dy A_x
        axis_terms(P, Lp, p, 2, 1, 1.0, l, m, G[0][0], G[0][1]);    // c = x: (a = z, b = y, +)
        axis_terms(P, Lp, p, 1, 2, -1.0, l, m, G[0][0], G[0][1]);   //        (a = y, b = z, -)
        axis_terms(P, Lp, p, 0, 2, 1.0, l, m, G[1][0], G[1][1]);    // c = y: (a = x, b = z, +)
        axis_terms(P, Lp, p, 2, 0, -1.0, l, m, G[1][0], G[1][1]);   //        (a = z, b = x, -)
        axis_terms(P, Lp, p, 1, 0, 1.0, l, m, G[2][0], G[2][1]);    // c = z: (a = y, b = x, +)
        axis_terms(P, Lp, p, 0, 1, -1.0, l, m, G[2][0], G[2][1]);   //        (a = x, b = y, -)
        double s = sgnl(l) * s2;
        for (int c = 0; c < 3; c++) {
            if (m == 0) out[c * Lp + lmi(l, 0)] = (T)(s * G[c][0]);
            else {
                out[c * Lp + lmi(l, m)] = (T)(s * SQRT2_A * G[c][0]);
                out[c * Lp + lmi(l, -m)] = (T)(s * SQRT2_A * G[c][1]);
            }
        }
    }
    const int L = (p + 1) * (p + 1);
    for (int e = threadIdx.x; e < 3 * Lp; e += blockDim.x)
        if (e % Lp >= L) out[e] = (T)0;
}

// dL/dq_j = sum_lm mu(l,m) R_lm(u_j): one wave per useful source leaf, one source per lane
template <class T>
__global__ __launch_bounds__(64) void k_p2m_adj(const int* __restrict__ leaves, const double* __restrict__ center,
                                                const int* __restrict__ begin, const int* __restrict__ end,
                                                const int* __restrict__ level, ScaleTable sig,
                                                const double* __restrict__ sxyz, const T* __restrict__ mu, int p,
                                                int Lp, double* __restrict__ gq) {
    extern __shared__ double al[];   // [3][NT][2] complex alpha'
    const int leaf = leaves[blockIdx.x];
    const int NT = (p + 1) * (p + 2) / 2;
    const T* M = mu + (size_t)leaf * 3 * Lp;
    for (int e = threadIdx.x; e < NT; e += 64) {
        int l = 0;
        while ((l + 1) * (l + 2) / 2 <= e) l++;
        int m = e - l * (l + 1) / 2;
        double s = sgnl(l);
        for (int c = 0; c < 3; c++) {
            double ar, ai;
            if (m == 0) { ar = s * (double)M[c * Lp + lmi(l, 0)]; ai = 0; }
            else { ar = s * SQRT2_A * (double)M[c * Lp + lmi(l, m)]; ai = -s * SQRT2_A * (double)M[c * Lp + lmi(l, -m)]; }
            al[(c * NT + e) * 2] = ar;
            al[(c * NT + e) * 2 + 1] = ai;
        }
    }
    __syncthreads();
    const double cx = center[3 * leaf], cy = center[3 * leaf + 1], cz = center[3 * leaf + 2];
    const double inv = 1.0 / sig.s[level[leaf]];
    for (int k = begin[leaf] + threadIdx.x; k < end[leaf]; k += 64) {
        double ux = (sxyz[3 * k] - cx) * inv, uy = (sxyz[3 * k + 1] - cy) * inv, uz = (sxyz[3 * k + 2] - cz) * inv;
        double r2 = ux * ux + uy * uy + uz * uz;
        double a0 = 0, a1 = 0, a2 = 0;
        double dr = 1.0, di = 0.0;
        for (int m = 0; m <= p; m++) {
            if (m > 0) {
                double s = 0.5 / m;
                double nr = (dr * ux - di * uy) * s, ni = (dr * uy + di * ux) * s;
                dr = nr; di = ni;
            }
            double p1r = dr, p1i = di, p2r = 0.0, p2i = 0.0;
            for (int l = m; l <= p; l++) {
                double vr, vi;
                if (l == m) { vr = dr; vi = di; }
                else {
                    double f = 1.0 / (double)(l * l - m * m);
                    vr = ((2.0 * l - 1) * uz * p1r - r2 * p2r) * f;
                    vi = ((2.0 * l - 1) * uz * p1i - r2 * p2i) * f;
                    p2r = p1r; p2i = p1i; p1r = vr; p1i = vi;
                }
                int e = l * (l + 1) / 2 + m;
                a0 += al[(0 * NT + e) * 2] * vr - al[(0 * NT + e) * 2 + 1] * vi;
                a1 += al[(1 * NT + e) * 2] * vr - al[(1 * NT + e) * 2 + 1] * vi;
                a2 += al[(2 * NT + e) * 2] * vr - al[(2 * NT + e) * 2 + 1] * vi;
            }
        }
        gq[3 * k] = a0;
        gq[3 * k + 1] = a1;
        gq[3 * k + 2] = a2;
    }
}

// near field transposed: per target leaf, targets staged in LDS; lanes walk the sources
// of each close node and add their gradient with one FP64 atomic per source and leaf
template <class R>
__global__ __launch_bounds__(64) void k_p2p_adj(const int* __restrict__ leaves, const double* __restrict__ tcenter,
                                                const int* __restrict__ tbegin, const int* __restrict__ tend,
                                                const int* __restrict__ close_begin, const int* __restrict__ close_end,
                                                const int* __restrict__ close_s, const int* __restrict__ sbegin,
                                                const int* __restrict__ send, const double* __restrict__ sxyz,
                                                const double* __restrict__ txyz, const double* __restrict__ g,
                                                double* __restrict__ gq) {
    __shared__ R tx[64], ty[64], tz[64], g0[64], g1[64], g2[64];
    const int leaf = leaves[blockIdx.x];
    const int lane = threadIdx.x;
    const double cx = tcenter[3 * leaf], cy = tcenter[3 * leaf + 1], cz = tcenter[3 * leaf + 2];
    const int cb = close_begin[leaf], ce = close_end[leaf];
    for (int k0 = tbegin[leaf]; k0 < tend[leaf]; k0 += 64) {
        int nk = min(64, tend[leaf] - k0);
        __syncthreads();
        if (lane < nk) {
            int k = k0 + lane;
            tx[lane] = (R)(txyz[3 * k] - cx); ty[lane] = (R)(txyz[3 * k + 1] - cy); tz[lane] = (R)(txyz[3 * k + 2] - cz);
            g0[lane] = (R)g[3 * k]; g1[lane] = (R)g[3 * k + 1]; g2[lane] = (R)g[3 * k + 2];
        }
        __syncthreads();
        for (int ci = cb; ci < ce; ci++) {
            int s = close_s[ci];
            for (int j = sbegin[s] + lane; j < send[s]; j += 64) {
                R yx = (R)(sxyz[3 * j] - cx), yy = (R)(sxyz[3 * j + 1] - cy), yz = (R)(sxyz[3 * j + 2] - cz);
                R a0 = 0, a1 = 0, a2 = 0;
                for (int i = 0; i < nk; i++) {
                    R dx = tx[i] - yx, dy = ty[i] - yy, dz = tz[i] - yz;
                    R r2 = dx * dx + dy * dy + dz * dz;
                    R ri = r2 > (R)0 ? (R)1 / sqrt(r2) : (R)0;
                    R ri3 = ri * ri * ri;
                    // (d x g) / r^3
                    a0 += (dy * g2[i] - dz * g1[i]) * ri3;
                    a1 += (dz * g0[i] - dx * g2[i]) * ri3;
                    a2 += (dx * g1[i] - dy * g0[i]) * ri3;
                }
                unsafeAtomicAdd(&gq[3 * j], (double)a0);
                unsafeAtomicAdd(&gq[3 * j + 1], (double)a1);
                unsafeAtomicAdd(&gq[3 * j + 2], (double)a2);
            }
        }
    }
}

__global__ void k_weights(const int* __restrict__ slots, int n, const double* __restrict__ sJ,
                          const double* __restrict__ gq, double* __restrict__ dLdw) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    dLdw[slots[k]] = sJ[3 * k] * gq[3 * k] + sJ[3 * k + 1] * gq[3 * k + 1] + sJ[3 * k + 2] * gq[3 * k + 2];
}


void run_adjoint(AdjointState& S, const AdjointArgs& a, const double* dLdB, double* dLdw, hipStream_t st) {
    // the transposed passes run on class_gemm, whose tiles hold at most 64 columns
    auto check_width = [](const PassView& v) {
        if (v.ntiles && v.width > 64) throw Error(1, "adjoint pass planned with tiles wider than 64 columns");
    };
    for (const PassView& v : a.m2m_levels) check_width(v);
    for (const PassView& v : a.l2l_levels) check_width(v);
    check_width(a.m2l);
    const TreeDev& Ts = *a.S;
    const TreeDev& Tt = *a.T;
    const bool f64 = a.prec == Precision::F64;
    const size_t es = f64 ? sizeof(double) : sizeof(float);
    const int nts = Tt.num_slots, nss = Ts.num_slots;
    double* g = S.g.get<double>(3 * (size_t)std::max(nts, 1));
    k_gather_g<<<blocks(nts), 256, 0, st>>>(Tt.slots.as<int>(), nts, dLdB, a.scale, g);
    // L2P^T
    size_t nT = (size_t)Tt.num_nodes * 3 * a.Lp, nS = (size_t)Ts.num_nodes * 3 * a.Lp;
    double* P = S.Gexp.get<double>(nT);
    void* lam = S.lam.reserve(nT * es);
    void* mu = S.mu.reserve(nS * es);
    FMM_HIP(hipMemsetAsync(lam, 0, nT * es, st));
    FMM_HIP(hipMemsetAsync(mu, 0, nS * es, st));
    if (a.n_tgt_leaves) {
        launch_p2m(true, a.tgt_leaves, a.n_tgt_leaves, Tt, a.sigT, a.txyz, g, a.p - 1, a.Lp, P, st);
        if (f64)
            k_l2p_adj<double><<<a.n_tgt_leaves, 64, 0, st>>>(a.tgt_leaves, Tt.level.as<int>(), a.sigT, P, a.p, a.Lp,
                                                            (double*)lam);
        else
            k_l2p_adj<float><<<a.n_tgt_leaves, 64, 0, st>>>(a.tgt_leaves, Tt.level.as<int>(), a.sigT, P, a.p, a.Lp,
                                                           (float*)lam);
    }
    // L2L^T bottom-up (forward pairs parent -> child; here child -> parent, atomic)
    for (int k = Tt.num_levels - 1; k >= 1; k--) {
        const PassView& v = a.l2l_levels[k];
        if (v.ntiles)
            class_gemm(a.prec, a.down_ops, a.down_ops_lo, v.tiles, v.ntiles, v.dst, v.src, lam, lam, a.Lp, EPI_ATOMIC,
                       st, true, 1, TRI_LOWER);
    }
    // M2L^T (forward far_s -> far_t; here far_t -> far_s): one slab per (pair, component), then
    // each source node sums its pairs' slabs in pair order
    if (a.m2l.ntiles) {
        const int npairs = a.lists->num_far;
        build_pair_index(a.m2l.dst, npairs, Ts.num_nodes, S.m2l_ix, st);
        void* slab = S.slab.reserve((size_t)3 * npairs * a.Lp * es);
        class_gemm(a.prec, a.down_ops, a.down_ops_lo, a.m2l.tiles, a.m2l.ntiles, a.m2l.src, a.m2l.dst, lam, slab,
                   a.Lp, EPI_STORE, st, true);
        reduce_pairs(f64, S.m2l_ix, slab, a.Lp, a.Lp, mu, st);
    }
    // M2M^T top-down (forward child -> parent; here parent -> child, one write per child)
    for (int k = 1; k < Ts.num_levels; k++) {
        const PassView& v = a.m2m_levels[k];
        if (v.ntiles)
            class_gemm(a.prec, a.up_ops, a.up_ops_lo, v.tiles, v.ntiles, v.dst, v.src, mu, mu, a.Lp, EPI_ADD, st, true, 1,
                       TRI_UPPER);
    }
    // P2M^T
    double* gq = S.gq.get<double>(3 * (size_t)std::max(nss, 1));
    FMM_HIP(hipMemsetAsync(gq, 0, 3 * (size_t)std::max(nss, 1) * sizeof(double), st));
    if (a.n_src_leaves) {
        int NT = (a.p + 1) * (a.p + 2) / 2;
        size_t lds = (size_t)6 * NT * sizeof(double);
        if (f64)
            k_p2m_adj<double><<<a.n_src_leaves, 64, lds, st>>>(a.src_leaves, Ts.center.as<double>(), Ts.begin.as<int>(),
                                                              Ts.end.as<int>(), Ts.level.as<int>(), a.sigS, a.sxyz,
                                                              (const double*)mu, a.p, a.Lp, gq);
        else
            k_p2m_adj<float><<<a.n_src_leaves, 64, lds, st>>>(a.src_leaves, Ts.center.as<double>(), Ts.begin.as<int>(),
                                                             Ts.end.as<int>(), Ts.level.as<int>(), a.sigS, a.sxyz,
                                                             (const float*)mu, a.p, a.Lp, gq);
    }
    // P2P^T
    if (a.near_field && a.n_tgt_leaves) {
        const Lists& L = *a.lists;
        if (f64)
            k_p2p_adj<double><<<a.n_tgt_leaves, 64, 0, st>>>(a.tgt_leaves, Tt.center.as<double>(), Tt.begin.as<int>(),
                                                            Tt.end.as<int>(), L.close_begin.as<int>(),
                                                            L.close_end.as<int>(), L.close_s.as<int>(),
                                                            Ts.begin.as<int>(), Ts.end.as<int>(), a.sxyz, a.txyz, g, gq);
        else
            k_p2p_adj<float><<<a.n_tgt_leaves, 64, 0, st>>>(a.tgt_leaves, Tt.center.as<double>(), Tt.begin.as<int>(),
                                                           Tt.end.as<int>(), L.close_begin.as<int>(),
                                                           L.close_end.as<int>(), L.close_s.as<int>(),
                                                           Ts.begin.as<int>(), Ts.end.as<int>(), a.sxyz, a.txyz, g, gq);
    }
    // dL/dw = J . dL/dq (input order; sources outside the root keep 0)
    FMM_HIP(hipMemsetAsync(dLdw, 0, (size_t)std::max(a.ns, 1) * sizeof(double), st));
    k_weights<<<blocks(nss), 256, 0, st>>>(Ts.slots.as<int>(), nss, a.sJ, gq, dLdw);
    FMM_HIP(hipGetLastError());
}

}  // namespace fmm

// FP64 translation-operator tables, generated on the GPU from solid-harmonic
// recurrences.  Each operator is the real-basis matrix T (row = output coefficient,
// column = input coefficient, padded to Lp) of one of the reference's complex
// translators, conjugated with the real<->complex maps of harmonics.cpp:281-320:
//   M2M  multipole_translator.cpp:3-32   out(l,m) += R(l-a, m-b) f1(m, m-b) in(a,b)
//   M2L  multipole_translator.cpp:148-176 out(l,m) += I(l+a, m-b) f2(m, b, a) in(a,b)
//   L2L  multipole_translator.cpp:201-232 out(l,m) += R(a-l, m-b) f3(m, b, a, l) in(a,b)
// and written for scaled expansions (multipoles divided by sigma^degree, locals
// multiplied by sigma^(degree+1)), which makes every operator of the shared frame
// independent of the level and keeps FP32/BF16 magnitudes uniform across degrees.
#include "fmm_tables.hpp"

namespace fmm {

struct cplx {
    double re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }

__device__ __forceinline__ int tri(int n, int k) { return n * (n + 1) / 2 + k; }   // (n, k>=0) packed
__device__ __forceinline__ double sgn_pow(int e) { return (e & 1) ? -1.0 : 1.0; }

// complex reference harmonic X(n, k) for signed k (X(n,-k) = conj X(n,k))
__device__ __forceinline__ cplx harm(const cplx* X, int n, int k) {
    if (k >= 0) return X[tri(n, k)];
    cplx v = X[tri(n, -k)];
    return {v.re, -v.im};
}

// kernel K(l, m; a, b) of the complex translator (see header comment); zero if absent
__device__ __forceinline__ cplx kernel_entry(int kind, const cplx* X, int pI, int l, int m, int a, int b) {
    int n, k = m - b;
    double f;
    if (kind == OPK_M2M) {
        n = l - a;
        if (n < 0 || k < -n || k > n) return {0, 0};
        f = sgn_pow(abs((abs(m) - abs(k) - abs(m - k)) / 2));
    } else if (kind == OPK_M2L) {
        n = l + a;
        if (n > pI || k < -n || k > n) return {0, 0};
        f = sgn_pow(abs((abs(m - b) - abs(m) - abs(b)) / 2)) * sgn_pow(a);
    } else {
        n = a - l;
        if (n < 0 || k < -n || k > n) return {0, 0};
        f = sgn_pow(abs((abs(b) - abs(m - b) - abs(m)) / 2)) * sgn_pow(a + l);
    }
    return cscale(harm(X, n, k), f);
}

__device__ __forceinline__ void lm_of(int i, int& l, int& m) {
    l = (int)sqrt((double)i);
    while ((l + 1) * (l + 1) <= i) l++;
    while (l * l > i) l--;
    m = i - l * l - l;
}

// One workgroup per operator.
__global__ void k_make_ops(const OpSpec* __restrict__ specs, int p, int pI_full, int Lp, double* __restrict__ out) {
    extern __shared__ cplx X[];   // packed (n, k>=0), n <= pH
    const OpSpec sp = specs[blockIdx.x];
    const int kind = sp.kind;
    const int pH = (kind == OPK_M2L) ? (sp.m2l_full ? pI_full : p) : p;
    const int pI = pH;
    double su = (kind == OPK_M2M) ? sp.sig_in : sp.sig_out;
    double ux = sp.t[0] / su, uy = sp.t[1] / su, uz = sp.t[2] / su;
    double r2 = ux * ux + uy * uy + uz * uz;
    // column k of the harmonic table, one thread per k
    for (int k = threadIdx.x; k <= pH; k += blockDim.x) {
        cplx d;
        if (kind == OPK_M2L) {
            // Psi_k^k = (2k-1)!! (x+iy)^k / r^(2k+1)
            double inv = 1.0 / sqrt(r2);
            d = {inv, 0};
            for (int i = 1; i <= k; i++) d = cscale(cmul(d, {ux, uy}), (2.0 * i - 1) / r2);
        } else {
            // Phi_k^k = (x+iy)^k / (2^k k!)
            d = {1, 0};
            for (int i = 1; i <= k; i++) d = cscale(cmul(d, {ux, uy}), 1.0 / (2.0 * i));
        }
        cplx pm2 = {0, 0}, pm1 = d;
        X[tri(k, k)] = cscale(d, sgn_pow(k));
        for (int n = k + 1; n <= pH; n++) {
            cplx v;
            if (kind == OPK_M2L)
                v = cscale(cadd(cscale(pm1, (2.0 * n - 1) * uz), cscale(pm2, -(double)((n - 1) * (n - 1) - k * k))),
                           1.0 / r2);
            else
                v = cscale(cadd(cscale(pm1, (2.0 * n - 1) * uz), cscale(pm2, -r2)), 1.0 / (double)(n * n - k * k));
            X[tri(n, k)] = cscale(v, sgn_pow(n));
            pm2 = pm1;
            pm1 = v;
        }
    }
    __syncthreads();
    const int L = (p + 1) * (p + 1);
    const double c = 0.70710678118654752440;
    const double ratio = sp.sig_in / sp.sig_out;
    double* T = out + (size_t)blockIdx.x * Lp * Lp;
    for (int e = threadIdx.x; e < Lp * Lp; e += blockDim.x) {
        int i = e / Lp, j = e % Lp;
        double val = 0;
        if (i < L && j < L) {
            int l, mr, a, br;
            lm_of(i, l, mr);
            lm_of(j, a, br);
            // complex input v(b) for the real unit vector j
            int nb = 0;
            int bs[2];
            cplx vs[2];
            if (br == 0) { bs[0] = 0; vs[0] = {1, 0}; nb = 1; }
            else if (br > 0) { bs[0] = br; vs[0] = {c, 0}; bs[1] = -br; vs[1] = {c, 0}; nb = 2; }
            else { bs[0] = -br; vs[0] = {0, c}; bs[1] = br; vs[1] = {0, -c}; nb = 2; }
            auto out_c = [&](int m) {
                cplx s = {0, 0};
                for (int q = 0; q < nb; q++) s = cadd(s, cmul(kernel_entry(kind, X, pI, l, m, a, bs[q]), vs[q]));
                return s;
            };
            if (mr == 0) val = out_c(0).re;
            else if (mr > 0) { cplx o1 = out_c(mr), o2 = out_c(-mr); val = c * (o1.re + o2.re); }
            else { cplx o1 = out_c(-mr), o2 = out_c(mr); val = c * (o1.im - o2.im); }
            // scaling of the expansions
            if (kind == OPK_M2M) val *= pow(ratio, l);             // (sig_c / sig_p)^l
            else if (kind == OPK_L2L) val *= pow(1.0 / ratio, a + 1); // (sig_c / sig_p)^(a+1)
            else val *= pow(ratio, a);                             // (sig_s / sig_t)^a
        }
        T[e] = val;
    }
}

__global__ void k_ops_to_f32(const double* __restrict__ in, float* __restrict__ out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}
__global__ void k_ops_to_bf16x2(const double* __restrict__ in, __bf16* __restrict__ hi, __bf16* __restrict__ lo,
                                size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = (float)in[i];
    __bf16 h = (__bf16)v;
    hi[i] = h;
    lo[i] = (__bf16)(v - (float)h);
}

void make_operators(const OpSpec* d_specs, int nops, int p, int Lp, double* d_out, hipStream_t st) {
    if (nops == 0) return;
    int pI = 2 * p;
    size_t lds = (size_t)(pI + 1) * (pI + 2) / 2 * sizeof(cplx);
    k_make_ops<<<nops, 256, lds, st>>>(d_specs, p, pI, Lp, d_out);
    FMM_HIP(hipGetLastError());
}
void operators_to_f32(const double* in, float* out, size_t n, hipStream_t st) {
    if (n) k_ops_to_f32<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(in, out, n);
}
void operators_to_bf16x2(const double* in, __bf16* hi, __bf16* lo, size_t n, hipStream_t st) {
    if (n) k_ops_to_bf16x2<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(in, hi, lo, n);
}

}  // namespace fmm

namespace fmm {

// M2L operator specs of the far-pair classes, from each class's first pair
__global__ void k_m2l_specs(const int* __restrict__ class_begin, int nc, const int* __restrict__ far_t,
                            const int* __restrict__ far_s, const double* __restrict__ sc, const double* __restrict__ tc,
                            const int* __restrict__ slev, const int* __restrict__ tlev, ScaleTable sigS, ScaleTable sigT,
                            int m2l_full, OpSpec* __restrict__ out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    int i = class_begin[c];
    int t = far_t[i], s = far_s[i];
    OpSpec o;
    for (int a = 0; a < 3; a++) o.t[a] = sc[3 * s + a] - tc[3 * t + a];
    o.sig_in = sigS.s[slev[s]];
    o.sig_out = sigT.s[tlev[t]];
    o.kind = OPK_M2L;
    o.m2l_full = m2l_full;
    out[c] = o;
}

void make_m2l_specs(const int* class_begin, int nc, const int* far_t, const int* far_s, const double* sc,
                    const double* tc, const int* slev, const int* tlev, const ScaleTable& sigS, const ScaleTable& sigT,
                    int m2l_full, OpSpec* out, hipStream_t st) {
    if (nc) k_m2l_specs<<<(nc + 255) / 256, 256, 0, st>>>(class_begin, nc, far_t, far_s, sc, tc, slev, tlev, sigS,
                                                          sigT, m2l_full, out);
}

}  // namespace fmm

namespace fmm {

// one operator per (src, dst) node pair: t = c_src - c_dst (M2L: source - target,
// L2L: parent - child, M2M: child - parent)
__global__ void k_pair_specs(const int* __restrict__ src, const int* __restrict__ dst, int n,
                             const double* __restrict__ csrc, const double* __restrict__ cdst,
                             const int* __restrict__ lsrc, const int* __restrict__ ldst, ScaleTable sig_src,
                             ScaleTable sig_dst, int kind, int m2l_full, OpSpec* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = src[i], d = dst[i];
    OpSpec o;
    for (int a = 0; a < 3; a++) o.t[a] = csrc[3 * s + a] - cdst[3 * d + a];
    o.sig_in = sig_src.s[lsrc[s]];
    o.sig_out = sig_dst.s[ldst[d]];
    o.kind = kind;
    o.m2l_full = m2l_full;
    out[i] = o;
}

void make_pair_specs(const int* src, const int* dst, int n, const double* csrc, const double* cdst, const int* lsrc,
                     const int* ldst, const ScaleTable& sig_src, const ScaleTable& sig_dst, int kind, int m2l_full,
                     OpSpec* out, hipStream_t st) {
    if (n) k_pair_specs<<<(n + 255) / 256, 256, 0, st>>>(src, dst, n, csrc, cdst, lsrc, ldst, sig_src, sig_dst, kind,
                                                         m2l_full, out);
}

}  // namespace fmm

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
#include <algorithm>
#include <climits>

#include "fmm_tables.hpp"
#include "fmm_recip.hpp"

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

// nsplit workgroups per operator, each owning a contiguous range of output rows (so a launch
// of few operators still fills the GPU).  LDS: the harmonic table X(n, k>=0), the degree/order
// of every coefficient index and the per-degree scale factor of the operator kind.
// output: 0 = double, 1 = float, 2 = bf16 hi + bf16 lo (of the float value)
template <int MODE>
__global__ __launch_bounds__(256) void k_make_ops(const OpSpec* __restrict__ specs, int p, int pI_full, int Lp,
                                                  void* __restrict__ out, __bf16* __restrict__ out_lo, int nsplit) {
    extern __shared__ cplx X[];   // packed (n, k>=0), n <= pH
    const int opi = blockIdx.x / nsplit, slice = blockIdx.x - opi * nsplit;
    const OpSpec sp = specs[opi];
    const int kind = sp.kind;
    const int pH = (kind == OPK_M2L) ? (sp.m2l_full ? pI_full : p) : p;
    const int pI = pH;
    const int L = (p + 1) * (p + 1);
    short2* lmt = (short2*)(X + (pI_full + 1) * (pI_full + 2) / 2);   // (l, m) of coefficient i
    double* fac = (double*)(lmt + ((L + 1) & ~1));                     // scale factor per degree
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
            for (int i = 1; i <= k; i++) d = cscale(cmul(d, {ux, uy}), recip(2 * i));
        }
        cplx pm2 = {0, 0}, pm1 = d;
        X[tri(k, k)] = cscale(d, sgn_pow(k));
        for (int n = k + 1; n <= pH; n++) {
            cplx v;
            if (kind == OPK_M2L)
                v = cscale(cadd(cscale(pm1, (2.0 * n - 1) * uz), cscale(pm2, -(double)((n - 1) * (n - 1) - k * k))),
                           1.0 / r2);
            else
                v = cscale(cadd(cscale(pm1, (2.0 * n - 1) * uz), cscale(pm2, -r2)), recip(n * n - k * k));
            X[tri(n, k)] = cscale(v, sgn_pow(n));
            pm2 = pm1;
            pm1 = v;
        }
    }
    const double ratio = sp.sig_in / sp.sig_out;
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
        int l, m;
        lm_of(i, l, m);
        lmt[i] = make_short2((short)l, (short)m);
    }
    // scaling of the expansions: M2M (sig_c / sig_p)^l (output degree), L2L (sig_c / sig_p)^(a+1)
    // and M2L (sig_s / sig_t)^a (input degree a)
    for (int n = threadIdx.x; n <= p; n += blockDim.x)
        fac[n] = kind == OPK_M2M ? pow(ratio, n) : kind == OPK_L2L ? pow(1.0 / ratio, n + 1) : pow(ratio, n);
    __syncthreads();
    const double c = 0.70710678118654752440;
    const size_t base = (size_t)opi * Lp * Lp;
    // output rows of this workgroup: [r0, r1)
    const int rsl = (Lp + nsplit - 1) / nsplit;
    const int r0 = min(Lp, slice * rsl), r1 = min(Lp, r0 + rsl);
    const int c0 = min(r0, L), c1 = min(r1, L);   // computed rows (< L)
    // the entries (i < L, j < L) of rows [c0, c1): e = (i - c0) L + j, (i, j) advanced by the
    // stride instead of divided out (the row index may advance by several)
    const int LL = (c1 - c0) * L;
    int i = c0 + threadIdx.x / L, j = threadIdx.x - (threadIdx.x / L) * L;
    const int di = blockDim.x / L, dj = blockDim.x - di * L;
    for (int e0 = 0; e0 < LL; e0 += blockDim.x, i += di, j += dj) {
        if (j >= L) { j -= L; i++; }
        const int e = e0 + threadIdx.x;
        const bool in = e < LL;
        double val = 0;
        if (in) {
            const short2 oi = lmt[i], oj = lmt[j];
            const int l = oi.x, mr = oi.y, a = oj.x, br = oj.y;
            // complex input v(b) for the real unit vector j: one or two complex orders b0, b1
            const bool two = br != 0;
            const int b0 = br >= 0 ? br : -br, b1 = -b0;
            const cplx v0 = br == 0 ? cplx{1, 0} : br > 0 ? cplx{c, 0} : cplx{0, c};
            const cplx v1 = br > 0 ? cplx{c, 0} : cplx{0, -c};
            auto out_c = [&](int m) {
                cplx s = cmul(kernel_entry(kind, X, pI, l, m, a, b0), v0);
                if (two) s = cadd(s, cmul(kernel_entry(kind, X, pI, l, m, a, b1), v1));
                return s;
            };
            if (mr == 0) val = out_c(0).re;
            else if (mr > 0) { cplx o1 = out_c(mr), o2 = out_c(-mr); val = c * (o1.re + o2.re); }
            else { cplx o1 = out_c(-mr), o2 = out_c(mr); val = c * (o1.im - o2.im); }
            val *= fac[kind == OPK_M2M ? l : a];
        }
        const size_t o = base + (size_t)i * Lp + j;
        if (MODE == 0) {
            if (in) ((double*)out)[o] = val;
        } else if (MODE == 1) {
            if (in) ((float*)out)[o] = (float)val;
        } else {
            float v = (float)val;
            __bf16 h = (__bf16)v;
            if (in) {
                ((__bf16*)out)[o] = h;
                out_lo[o] = (__bf16)(v - (float)h);
            }
        }
    }
    // zero padding of rows [r0, r1): rows >= L whole, then columns L .. Lp-1 of rows < L
    const int np = Lp - L, z0 = max(r0, L);
    for (int e = z0 * Lp + threadIdx.x; e < r1 * Lp; e += blockDim.x) {
        if (MODE == 0) ((double*)out)[base + e] = 0.0;
        else if (MODE == 1) ((float*)out)[base + e] = 0.0f;
        else { ((__bf16*)out)[base + e] = (__bf16)0.0f; out_lo[base + e] = (__bf16)0.0f; }
    }
    for (int e = threadIdx.x; e < (c1 - c0) * np; e += blockDim.x) {
        const int r = c0 + e / np;
        const size_t o = base + (size_t)r * Lp + L + (e - (r - c0) * np);
        if (MODE == 0) ((double*)out)[o] = 0.0;
        else if (MODE == 1) ((float*)out)[o] = 0.0f;
        else { ((__bf16*)out)[o] = (__bf16)0.0f; out_lo[o] = (__bf16)0.0f; }
    }
}

constexpr int M2L_F16_CGROUPS = 2;   // column groups of 256 in the split (Kp <= 512)

// per M2L operator: the irregular harmonic table X(n, k >= 0), n <= pH, packed as in k_make_ops
// (nX entries), and the input-degree factors fac[a] = (sig_s / sig_t)^a.  Thread (op, k) runs
// column k of the recurrences; threads k <= p also give fac[k].
__global__ void k_m2l_harmonics(const OpSpec* __restrict__ specs, int nops, int p, int pI_full, int nX,
                                cplx* __restrict__ Xall, double* __restrict__ facall) {
    const int cols = pI_full + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nops * cols) return;
    const int opi = t / cols, k = t - opi * cols;
    const OpSpec sp = specs[opi];
    const int pH = sp.m2l_full ? pI_full : p;
    if (k <= p) facall[(size_t)opi * (p + 1) + k] = pow(sp.sig_in / sp.sig_out, k);
    if (k > pH) return;
    cplx* X = Xall + (size_t)opi * nX;
    const double ux = sp.t[0] / sp.sig_out, uy = sp.t[1] / sp.sig_out, uz = sp.t[2] / sp.sig_out;
    const double r2 = ux * ux + uy * uy + uz * uz;
    const double inv = 1.0 / sqrt(r2);
    cplx d = {inv, 0};   // Psi_k^k = (2k-1)!! (x+iy)^k / r^(2k+1)
    for (int i = 1; i <= k; i++) d = cscale(cmul(d, {ux, uy}), (2.0 * i - 1) / r2);
    cplx pm2 = {0, 0}, pm1 = d;
    X[tri(k, k)] = cscale(d, sgn_pow(k));
    for (int n = k + 1; n <= pH; n++) {
        cplx v = cscale(cadd(cscale(pm1, (2.0 * n - 1) * uz), cscale(pm2, -(double)((n - 1) * (n - 1) - k * k))),
                        1.0 / r2);
        X[tri(n, k)] = cscale(v, sgn_pow(n));
        pm2 = pm1;
        pm1 = v;
    }
}

// M2L operators as FP16 splits (F16Split), one item per (output degree l, order |m|) x (input
// degree a, order |b|): the complex kernel entries K(+-m, +-b) = f(+-m, +-b) X(l + a, +-m -+ b)
// take two harmonics, X(n, m - b) and X(n, m + b), and their conjugates, and give the up to four
// real entries of rows (l, +-m) and columns (a, +-b) with the same operations, in the same
// order, as the entry-wise kernel.  Workgroup s of an operator owns the row pairs (l, |m|) of
// packed index [s pps, (s+1) pps), staged in LDS as FP32 rows (scaled by fac[a] and 1 / dm_j),
// two per pair (+m, then -m; m = 0 leaves the second unused); each wave then takes whole rows:
// the row maximum by a shuffle reduction, the power-of-2 row scale and the hi + lo split into
// the [R][Kp] tables.  Rows L .. R-1 (zero, rinv 1) belong to workgroup 0.
__global__ __launch_bounds__(256) void k_make_m2l_f16(const OpSpec* __restrict__ specs, int p, int pI_full, int nX,
                                                      const cplx* __restrict__ Xall,
                                                      const double* __restrict__ facall, F16Split f16, int pps,
                                                      int nsplit) {
    extern __shared__ cplx X[];   // packed (n, k>=0), n <= pH
    const int opi = blockIdx.x / nsplit, slice = blockIdx.x - opi * nsplit;
    const int pH = specs[opi].m2l_full ? pI_full : p;
    const int L = (p + 1) * (p + 1), NP = (p + 1) * (p + 2) / 2;
    short2* tri_lm = (short2*)(X + nX);                    // (n, k) of packed index t (degree <= p)
    double* cf = (double*)(tri_lm + ((NP + 1) & ~1));      // column factors: fac[a] and 1 / dm_j
    double* idm = cf + (p + 1);
    float* stg = (float*)(idm + L);                        // the slice's rows, [row][L]
    {
        const cplx* Xg = Xall + (size_t)opi * nX;
        const int nXh = (pH + 1) * (pH + 2) / 2;
        for (int t = threadIdx.x; t < nXh; t += blockDim.x) X[t] = Xg[t];
    }
    for (int t = threadIdx.x; t < NP; t += blockDim.x) {
        int n = 0;
        while ((n + 1) * (n + 2) / 2 <= t) n++;
        tri_lm[t] = make_short2((short)n, (short)(t - n * (n + 1) / 2));
    }
    for (int n = threadIdx.x; n <= p; n += blockDim.x) cf[n] = facall[(size_t)opi * (p + 1) + n];
    for (int j = threadIdx.x; j < L; j += blockDim.x) idm[j] = (double)(1.0f / f16.dm[j]);   // exact: dm is 2^k
    __syncthreads();
    const double c = 0.70710678118654752440;
    const int rp0 = slice * pps, nrp = min(NP, rp0 + pps) - rp0;
    // items e = rp NP + q, (rp, q) advanced by the stride instead of divided out
    int rp = threadIdx.x / NP, q = threadIdx.x - (threadIdx.x / NP) * NP;
    const int drp = blockDim.x / NP, dq = blockDim.x - drp * NP;
    for (int e = threadIdx.x; e < nrp * NP; e += blockDim.x, rp += drp, q += dq) {
        if (q >= NP) { q -= NP; rp++; }
        const short2 lm = tri_lm[rp0 + rp], ab = tri_lm[q];
        const int l = lm.x, m = lm.y, a = ab.x, b = ab.y, n = l + a;
        // K(s1 m, s2 b), s = +-1: f(m', b') = (-1)^|(|m'-b'| - |m'| - |b'|) / 2| (-1)^a
        cplx kpp = {0, 0}, kpm = {0, 0}, kmp = {0, 0}, kmm = {0, 0};
        if (n <= pH) {
            const cplx A = harm(X, n, m - b), Bq = harm(X, n, m + b);
            const double sa = sgn_pow(a);
            kpp = cscale(A, sgn_pow(abs((abs(m - b) - m - b) / 2)) * sa);
            kpm = cscale(Bq, sgn_pow(abs((abs(m + b) - m - b) / 2)) * sa);
            kmp = cscale(cplx{Bq.re, -Bq.im}, sgn_pow(abs((abs(-m - b) - m - b) / 2)) * sa);
            kmm = cscale(cplx{A.re, -A.im}, sgn_pow(abs((abs(b - m) - m - b) / 2)) * sa);
        }
        const double fa = cf[a];
        const int rp_ = 2 * rp, rm_ = 2 * rp + 1;   // staged rows (l, +-m)
        const int jp = a * a + a + b, jm = a * a + a - b;   // columns (a, +-b)
        auto put = [&](int r, int j, double val) {
            val *= fa;
            stg[r * L + j] = (float)(val * idm[j]);
        };
        if (b == 0) {   // column (a, 0): out_c(m') = K(m', 0)
            if (m == 0) put(rp_, jp, kpp.re);
            else {
                put(rp_, jp, c * (kpp.re + kmp.re));
                put(rm_, jp, c * (kpp.im - kmp.im));
            }
        } else {
            // column (a, +b): out_c(m') = c K(m', b) + c K(m', -b)
            // column (a, -b): out_c(m') = (-c Im K(m', b) + c Im K(m', -b), c Re K(m', b) - c Re K(m', -b))
            const cplx opP = {c * kpp.re + c * kpm.re, c * kpp.im + c * kpm.im};       // m' = m, col +b
            const cplx opM = {-(c * kpp.im) + c * kpm.im, c * kpp.re + -(c * kpm.re)};  // m' = m, col -b
            if (m == 0) {
                put(rp_, jp, opP.re);
                put(rp_, jm, opM.re);
            } else {
                const cplx omP = {c * kmp.re + c * kmm.re, c * kmp.im + c * kmm.im};       // m' = -m, col +b
                const cplx omM = {-(c * kmp.im) + c * kmm.im, c * kmp.re + -(c * kmm.re)};  // m' = -m, col -b
                put(rp_, jp, c * (opP.re + omP.re));
                put(rm_, jp, c * (opP.im - omP.im));
                put(rp_, jm, c * (opM.re + omM.re));
                put(rm_, jm, c * (opM.im - omM.im));
            }
        }
    }
    __syncthreads();
    // rows: maximum, scale, split.  Lane g takes the columns 4g .. 4g + 3 of each group of 256.
    typedef _Float16 h4 __attribute__((ext_vector_type(4)));
    const int R = f16.R, Kp = f16.Kp, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t ob = (size_t)opi * R * Kp;
    const int nst = 2 * nrp, r_end = nst + (slice == 0 ? R - L : 0);   // workgroup 0 also zeroes rows L .. R-1
    for (int r = wave; r < r_end; r += 4) {
        const bool pad = r >= nst;
        int i = L + (r - nst);
        if (!pad) {
            const short2 lm = tri_lm[rp0 + (r >> 1)];
            if ((r & 1) && lm.y == 0) continue;   // m = 0: one row per pair (uniform over the wave)
            i = lm.x * lm.x + lm.x + ((r & 1) ? -lm.y : lm.y);
        }
        float v[M2L_F16_CGROUPS][4];
        float mx = 0.0f;
#pragma unroll
        for (int g = 0; g < M2L_F16_CGROUPS; g++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int j = 256 * g + 4 * lane + k;
                v[g][k] = (!pad && j < L) ? stg[r * L + j] : 0.0f;
                mx = fmaxf(mx, fabsf(v[g][k]));
            }
        for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
        const float rs = mx > 0.0f ? exp2f((float)(-(ilogbf(mx) + 1))) : 1.0f;
#pragma unroll
        for (int g = 0; g < M2L_F16_CGROUPS; g++) {
            const int j = 256 * g + 4 * lane;
            if (j >= Kp) continue;
            h4 hi, lo;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float x = v[g][k] * rs;
                const _Float16 h = (_Float16)x;
                hi[k] = h;
                lo[k] = (_Float16)(x - (float)h);
            }
            *(h4*)(f16.hi + ob + (size_t)i * Kp + j) = hi;
            *(h4*)(f16.lo + ob + (size_t)i * Kp + j) = lo;
        }
        if (lane == 0) f16.rinv[(size_t)opi * R + i] = 1.0f / rs;
    }
}

// row pairs per workgroup of k_make_m2l_f16: as many as the staging budget holds (two rows of
// L floats each), so that several workgroups share a CU
constexpr size_t M2L_F16_STAGE = 16 * 1024;   // p = 12: 12 pairs, 23 KB in all, 7 workgroups per CU

size_t m2l_f16_work_bytes(int nops, int p) {
    const int pI = 2 * p, nX = (pI + 1) * (pI + 2) / 2;
    return (size_t)nops * nX * sizeof(cplx) + (size_t)nops * (p + 1) * sizeof(double);
}

void make_m2l_f16(const OpSpec* d_specs, int nops, int p, const F16Split& f16, void* work, hipStream_t st) {
    if (nops == 0) return;
    const int L = (p + 1) * (p + 1), NP = (p + 1) * (p + 2) / 2, pI = 2 * p, nX = (pI + 1) * (pI + 2) / 2;
    if (f16.R < L || f16.Kp < L || f16.Kp % 4 || f16.Kp > 256 * M2L_F16_CGROUPS)
        throw Error(4, "FP16 split tables: bad row/column padding");
    cplx* Xall = (cplx*)work;
    double* facall = (double*)(Xall + (size_t)nops * nX);
    const int nthr = nops * (pI + 1);
    k_m2l_harmonics<<<(nthr + 255) / 256, 256, 0, st>>>(d_specs, nops, p, pI, nX, Xall, facall);
    FMM_HIP(hipGetLastError());
    const size_t per_pair = 2 * (size_t)L * sizeof(float);
    const int pps = (int)std::max<size_t>(1, std::min<size_t>(NP, M2L_F16_STAGE / per_pair));
    const int nsplit = (NP + pps - 1) / pps;
    const size_t lds = (size_t)nX * sizeof(cplx) + (size_t)((NP + 1) & ~1) * sizeof(short2) +
                       (size_t)(p + 1 + L) * sizeof(double) + per_pair * pps;
    if (lds > 64 * 1024) throw Error(4, "FP16 split tables: order too large for the LDS staging");
    k_make_m2l_f16<<<nops * nsplit, 256, lds, st>>>(d_specs, p, pI, nX, Xall, facall, f16, pps, nsplit);
    FMM_HIP(hipGetLastError());
}

// Per-node M2M: the reference's ComplexTranslation / RealTranslation / Layers algorithms
// (multipole_solver.cpp:141-339; the translator multipole_translator.cpp:3-32), which translate
// each child's expansion with the formula out(l, m) += sum R(l - a, m - b) f(m, m - b) in(a, b)
// evaluated from the child offset's harmonics, without operator tables.  One workgroup per
// parent: per non-empty child, the complex harmonics X of u = (c_child - c_parent) / sigma_child
// and the child's expansion in the complex basis go to LDS, then every thread owns output
// coefficients i = tid + 256 r, sums the complex formula for (l, +-m), converts to the real basis
// and scales by (sigma_child / sigma_parent)^l; the sums over the children stay in registers and
// are added to the parent once.
template <class T>
__global__ __launch_bounds__(256) void k_m2m_nodes(const int* __restrict__ parents, const int* __restrict__ first_child,
                                                   const int* __restrict__ npts, const double* __restrict__ center,
                                                   const int* __restrict__ level, ScaleTable sig, int p, int Lp,
                                                   T* __restrict__ M) {
    extern __shared__ cplx sh[];
    constexpr int RMAX = (MAX_ORDER + 1) * (MAX_ORDER + 1) / 256 + 1;
    const int L = (p + 1) * (p + 1), NX = (p + 1) * (p + 2) / 2;
    cplx* X = sh;            // packed (n, k >= 0), n <= p
    cplx* C = sh + NX;       // [3][L] complex input, index l^2 + l + m
    const int parent = parents[blockIdx.x], fc = first_child[parent];
    const double c = 0.70710678118654752440;
    double acc[RMAX][3];
    for (int r = 0; r < RMAX; r++) acc[r][0] = acc[r][1] = acc[r][2] = 0.0;
    const double sp = sig.s[level[parent]];
    for (int o = 0; o < 8; o++) {
        const int ch = fc + o;
        if (first_child[ch] < 0 && npts[ch] == 0) continue;   // uniform over the workgroup
        const double sc = sig.s[level[ch]];
        __syncthreads();   // the previous child's tables are no longer read
        const double ux = (center[3 * ch] - center[3 * parent]) / sc, uy = (center[3 * ch + 1] - center[3 * parent + 1]) / sc,
                     uz = (center[3 * ch + 2] - center[3 * parent + 2]) / sc;
        const double r2 = ux * ux + uy * uy + uz * uz;
        for (int k = threadIdx.x; k <= p; k += blockDim.x) {   // Phi_k^k = (x+iy)^k / (2^k k!), column k
            cplx d = {1, 0};
            for (int i = 1; i <= k; i++) d = cscale(cmul(d, {ux, uy}), recip(2 * i));
            cplx pm2 = {0, 0}, pm1 = d;
            X[tri(k, k)] = cscale(d, sgn_pow(k));
            for (int n = k + 1; n <= p; n++) {
                cplx v = cscale(cadd(cscale(pm1, (2.0 * n - 1) * uz), cscale(pm2, -r2)), recip(n * n - k * k));
                X[tri(n, k)] = cscale(v, sgn_pow(n));
                pm2 = pm1;
                pm1 = v;
            }
        }
        const T* Mc = M + (size_t)ch * 3 * Lp;
        for (int e = threadIdx.x; e < 3 * L; e += blockDim.x) {   // real -> complex (harmonics.cpp:281-302)
            const int comp = e / L, i = e - comp * L;
            int a, b;
            lm_of(i, a, b);
            const double re = (double)Mc[comp * Lp + a * a + a + abs(b)];
            const double im = b == 0 ? 0.0 : (double)Mc[comp * Lp + a * a + a - abs(b)];
            C[e] = b == 0 ? cplx{re, 0.0} : b > 0 ? cplx{c * re, c * im} : cplx{c * re, -c * im};
        }
        __syncthreads();
        const double ratio = sc / sp;
        for (int r = 0; r < RMAX; r++) {
            const int i = threadIdx.x + 256 * r;
            if (i >= L) break;
            int l, m;
            lm_of(i, l, m);
            const int am = abs(m);
            const double f = pow(ratio, l);
            for (int comp = 0; comp < 3; comp++) {
                const cplx* Cc = C + comp * L;
                cplx op = {0, 0}, om = {0, 0};   // complex outputs at (l, |m|) and (l, -|m|)
                for (int a = 0; a <= l; a++)
                    for (int b = -a; b <= a; b++) {
                        const cplx in = Cc[a * a + a + b];
                        op = cadd(op, cmul(kernel_entry(OPK_M2M, X, p, l, am, a, b), in));
                        if (am) om = cadd(om, cmul(kernel_entry(OPK_M2M, X, p, l, -am, a, b), in));
                    }
                // complex -> real (harmonics.cpp:304-320)
                const double v = m == 0 ? op.re : m > 0 ? c * (op.re + om.re) : c * (op.im - om.im);
                acc[r][comp] += v * f;
            }
        }
    }
    T* Mp = M + (size_t)parent * 3 * Lp;
    for (int r = 0; r < RMAX; r++) {
        const int i = threadIdx.x + 256 * r;
        if (i >= L) break;
        for (int comp = 0; comp < 3; comp++) Mp[comp * Lp + i] += (T)acc[r][comp];
    }
}

void m2m_per_node(bool f64, const int* parents, int nparents, const TreeDev& S, const ScaleTable& sig, int p, int Lp,
                  void* M, hipStream_t st) {
    if (nparents <= 0) return;
    if (p > MAX_ORDER) throw Error(4, "expansion order above 20");
    const int L = (p + 1) * (p + 1), NX = (p + 1) * (p + 2) / 2;
    const size_t lds = (size_t)(NX + 3 * L) * sizeof(cplx);
    if (f64)
        k_m2m_nodes<double><<<nparents, 256, lds, st>>>(parents, S.first_child.as<int>(), S.npts.as<int>(),
                                                        S.center.as<double>(), S.level.as<int>(), sig, p, Lp,
                                                        (double*)M);
    else
        k_m2m_nodes<float><<<nparents, 256, lds, st>>>(parents, S.first_child.as<int>(), S.npts.as<int>(),
                                                       S.center.as<double>(), S.level.as<int>(), sig, p, Lp,
                                                       (float*)M);
    FMM_HIP(hipGetLastError());
}

void make_operators(const OpSpec* d_specs, int nops, int p, int Lp, int out_mode, void* d_out, __bf16* d_out_lo,
                    hipStream_t st) {
    if (nops == 0) return;
    int pI = 2 * p, L = (p + 1) * (p + 1);
    size_t lds = (size_t)(pI + 1) * (pI + 2) / 2 * sizeof(cplx) + (size_t)((L + 1) & ~1) * sizeof(short2) +
                 (size_t)(p + 1) * sizeof(double);
    // about 2048 workgroups in all, at least 8 rows each
    const int nsplit = std::max(1, std::min((Lp + 7) / 8, 2048 / nops));
    const int grid = nops * nsplit;
    if (out_mode == 0) k_make_ops<0><<<grid, 256, lds, st>>>(d_specs, p, pI, Lp, d_out, nullptr, nsplit);
    else if (out_mode == 1) k_make_ops<1><<<grid, 256, lds, st>>>(d_specs, p, pI, Lp, d_out, nullptr, nsplit);
    else k_make_ops<2><<<grid, 256, lds, st>>>(d_specs, p, pI, Lp, d_out, d_out_lo, nsplit);
    FMM_HIP(hipGetLastError());
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

// M2M operator specs per (child level k, orientation o), as the reference builds them
// (multipole_solver.cpp:472-534): from the first non-empty child of orientation o at level k (the
// lowest node id, BFS order) and its parent, so the translation is the float-rounded offset of
// that pair; an orientation with no such child keeps the nominal offset (no node applies it).
// One workgroup per child level.
struct LevelRanges {
    int begin[MAX_LEVELS + 2];
};
__global__ __launch_bounds__(64) void k_m2m_specs(LevelRanges lv, int nlev, const int* __restrict__ first_child,
                                                  const int* __restrict__ npts, const int* __restrict__ parent,
                                                  const double* __restrict__ center, ScaleTable sig,
                                                  const double* __restrict__ nominal, OpSpec* __restrict__ out) {
    // One wave per child level.  Children are numbered breadth first, parent by parent, so the
    // first parent (in level order) holding a non-empty child of orientation o holds the
    // smallest such child: the wave walks the parents 64 at a time and stops once all eight
    // orientations are found (usually in the first chunk).
    const int k = blockIdx.x + 1, lane = threadIdx.x;
    int best = INT_MAX;   // lane o < 8: the child of orientation o
    unsigned found = 0;   // bit o: orientation o found (uniform over the wave)
    for (int n0 = lv.begin[k - 1]; n0 < lv.begin[k] && found != 0xffu; n0 += 64) {
        const int n = n0 + lane;
        const int fc = n < lv.begin[k] ? first_child[n] : -1;
        for (int o = 0; o < 8; o++) {
            if ((found >> o) & 1u) continue;
            const bool ok = fc >= 0 && (first_child[fc + o] >= 0 || npts[fc + o] > 0);
            const unsigned long long b = __ballot(ok);
            if (b == 0) continue;
            const int ch = __shfl(fc + o, __ffsll(b) - 1);
            if (lane == o) best = ch;
            found |= 1u << o;
        }
    }
    if (lane < 8) {
        const int o = lane, ch = best;
        OpSpec sp;
        for (int a = 0; a < 3; a++)
            sp.t[a] = ch == INT_MAX ? nominal[((k - 1) * 8 + o) * 3 + a]
                                    : center[3 * ch + a] - center[3 * parent[ch] + a];   // child - parent
        sp.sig_in = sig.s[k];
        sp.sig_out = sig.s[k - 1];
        sp.kind = OPK_M2M;
        sp.m2l_full = 0;
        out[(k - 1) * 8 + o] = sp;
    }
}

void make_m2m_specs(const TreeDev& S, const ScaleTable& sig, const double* d_nominal, OpSpec* out, hipStream_t st) {
    if (S.num_levels < 2) return;
    LevelRanges lv;
    for (int k = 0; k < MAX_LEVELS + 2; k++) lv.begin[k] = S.level_begin[k];
    k_m2m_specs<<<S.num_levels - 1, 64, 0, st>>>(lv, S.num_levels, S.first_child.as<int>(), S.npts.as<int>(),
                                                   S.parent.as<int>(), S.center.as<double>(), sig, d_nominal, out);
    FMM_HIP(hipGetLastError());
}

void make_pair_specs(const int* src, const int* dst, int n, const double* csrc, const double* cdst, const int* lsrc,
                     const int* ldst, const ScaleTable& sig_src, const ScaleTable& sig_dst, int kind, int m2l_full,
                     OpSpec* out, hipStream_t st) {
    if (n) k_pair_specs<<<(n + 255) / 256, 256, 0, st>>>(src, dst, n, csrc, cdst, lsrc, ldst, sig_src, sig_dst, kind,
                                                         m2l_full, out);
}

}  // namespace fmm

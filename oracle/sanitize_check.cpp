// sanitize_check.cpp -- TEST INFRASTRUCTURE: a driver for the oracle under AddressSanitizer and
// UndefinedBehaviorSanitizer (built and run by tests/test_oracle_sanitized.py, never shipped).
//
// It includes the oracle's single source file and runs every entry point the tests use on small
// seeded inputs -- harmonics, translators, both tree kinds with a depth cap, the interaction
// lists, the whole FMM in the reference's flags and in the corrected ones (with and without
// per-source current directions), the treecode, BEM, and the direct sums -- so that an
// out-of-bounds access, a use after free or undefined arithmetic anywhere on those paths ends
// the run with the sanitizer's report.  Exit status 0 means every path ran clean; the numbers
// themselves are checked by the other oracle tests.
#include "fmm_oracle.cpp"

#include <cstdio>
#include <random>

namespace {

std::vector<double> points(std::mt19937_64& g, size_t n, double lo, double hi) {
    std::uniform_real_distribution<double> u(lo, hi);
    std::vector<double> v(3 * n);
    for (auto& x : v) x = u(g);
    return v;
}

orc_config config(int p, int cap, int flags) {
    orc_config c{};
    c.p = p;
    c.cap_s = cap;
    c.cap_t = cap;
    c.flags = flags;
    c.target_root_mode = 0;   // bounding box of the targets
    c.max_depth = 21;
    c.mac_scale = 1.0;
    c.current = 1.0;
    c.fd_eps = 0.0;           // the reference's step
    c.problem = 0;
    for (int a = 0; a < 3; a++) { c.source_root_c[a] = 0.0; c.source_root_h[a] = 3.0; }
    return c;
}

}  // namespace

int main() {
    std::mt19937_64 g(20261202);
    const int corrected = SECTORAL_FIX | FACT_FIX | ANALYTIC_CURL | P2P | M2L_FULL;
    // harmonics and translators at the orders the tests use
    for (int p : {4, 8, 12}) {
        const double v[3] = {0.3, -0.2, 0.5}, t[3] = {0.25, 0.25, -0.25};
        std::vector<double> h(2 * orc_nlen(p)), Tm((size_t)(p + 1) * (p + 1) * (p + 1) * (p + 1));
        orc_solid_harmonics(p, v, 1, 0, h.data());
        orc_solid_harmonics(p, v, 0, corrected, h.data());
        orc_solid_harmonics_cart(p, v, 1, h.data());
        std::vector<double> grad(3 * (p + 1) * (p + 1));
        orc_regular_grad(p, v, grad.data());
        for (int kind = 0; kind < 3; kind++) orc_operator(kind, p, t, corrected, Tm.data());
        orc_m2m_matrix_operator(p, t, 0, Tm.data());
    }
    // trees (a cluster deep enough to reach a depth cap), lists
    const size_t ns = 3000, nt = 700;
    std::vector<double> src = points(g, ns, -1.0, 1.0), tgt = points(g, nt, -1.5, 1.5);
    for (size_t i = 0; i < 200; i++)
        for (int a = 0; a < 3; a++) src[3 * i + a] = 0.2 + 1e-7 * (double)(i % 7) * (a + 1);
    std::vector<double> w(ns, 1.0 / ns), J(3 * ns);
    for (size_t i = 0; i < ns; i++) { J[3 * i] = 0.0; J[3 * i + 1] = 0.6; J[3 * i + 2] = 0.8; }
    const double rc[3] = {0, 0, 0}, rh[3] = {3, 3, 3};
    void* S = orc_tree_build(ns, src.data(), 1, rc, rh, 32, 12);
    void* T = orc_tree_build(nt, tgt.data(), 0, rc, rh, 33, 21);
    void* Lh = orc_lists_build(S, T, 1.0);
    size_t nfar = 0, nclose = 0;
    orc_lists_counts(Lh, &nfar, &nclose);
    std::vector<int> ft(nfar), fs(nfar), ct(nclose), cs(nclose);
    orc_lists_get(Lh, S, T, ft.data(), fs.data(), ct.data(), cs.data());
    orc_lists_free(Lh);
    orc_tree_free(S);
    orc_tree_free(T);
    // the FMM in both flag sets, with and without current directions; treecode
    for (int flags : {0, corrected}) {
        for (const double* jp : {(const double*)nullptr, (const double*)J.data()}) {
            orc_config c = config(8, 32, flags);
            void* R = orc_fmm_run(&c, ns, src.data(), w.data(), jp, nt, tgt.data());
            const int nslots = orc_fmm_num_slots(R);
            std::vector<int> slots(nslots);
            std::vector<double> B(3 * nslots), Bf(3 * nslots), Bn(3 * nslots), A(3 * nslots), Af(3 * nslots),
                An(3 * nslots);
            orc_fmm_get(R, slots.data(), B.data(), Bf.data(), Bn.data());
            orc_fmm_get_A(R, A.data(), Af.data(), An.data());
            const double x[6] = {10, 10, 10, 9.5, 10.5, 10.2};
            double tb[6], ta[6];
            orc_treecode(R, 0, 2, x, tb, ta);
            std::vector<double> loc(3 * (9 * 9)), mul(3 * (9 * 9));
            orc_fmm_local(R, 0, loc.data());
            orc_fmm_multipole(R, 0, mul.data());
            orc_fmm_free(R);
        }
    }
    // BEM (its own source root) and the direct sums
    {
        std::vector<double> Bv = points(g, ns, -1.0, 1.0), nrm = points(g, ns, -1.0, 1.0);
        orc_config c = config(6, 32, corrected);
        c.problem = 1;
        c.source_root_h[0] = c.source_root_h[1] = 1.1;
        c.source_root_h[2] = 1.6;
        void* R = orc_fmm_run_bem(&c, ns, src.data(), w.data(), Bv.data(), nrm.data(), nt, tgt.data());
        orc_fmm_free(R);
        std::vector<double> out(3 * nt);
        orc_direct_bem(ns, src.data(), w.data(), Bv.data(), nrm.data(), nt, tgt.data(), out.data());
    }
    std::vector<double> B(3 * nt), A(3 * nt), gw(ns);
    orc_direct_B(ns, src.data(), w.data(), J.data(), nt, tgt.data(), 1.0, B.data());
    orc_direct_A(ns, src.data(), w.data(), nullptr, nt, tgt.data(), 1.0, A.data());
    orc_direct_grad_w(ns, src.data(), w.data(), J.data(), nt, tgt.data(), 1.0, B.data(), gw.data());
    std::printf("sanitized oracle run: %zu far and %zu close pairs, every entry point ran\n", nfar, nclose);
    return 0;
}

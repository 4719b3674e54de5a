// Drives the fused evaluation (two host threads, pinned staging buffers), the adjoint and the
// allocation-free direct sum through the C ABI at p = 12, 14, 16 in F16X3 and FP32 -- built by
// build_asan.py against the library with AddressSanitizer on its host code, so that any host-side
// memory error on those paths is reported.  Prints the rel-L2 of each run against the direct sum.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <random>
#include <vector>
#include "fmm_capi.h"
#define CK(x) do { int rc_ = (x); if (rc_) { std::printf("FAIL %s rc %d: %s\n", #x, rc_, fmm_last_error(c)); std::_Exit(1); } } while (0)
int main() {
    std::setvbuf(stdout, nullptr, _IONBF, 0);
    for (int order : {14, 16, 12}) {
        for (int prec : {FMM_PREC_F16X3, FMM_PREC_FP32}) {
            fmm_config cfg;
            fmm_config_init(&cfg);
            cfg.order = order;
            cfg.precision = prec;
            cfg.frame = FMM_FRAME_SHARED;
            cfg.source_leaf_capacity = 64;
            cfg.target_leaf_capacity = 64;
            fmm_ctx* c = nullptr;
            if (fmm_create(&c, &cfg)) { std::printf("create failed\n"); return 1; }
            std::mt19937_64 g(11 + order + prec);
            std::uniform_real_distribution<double> U(0, 1);
            const size_t n = 60000, m = 20000;
            std::vector<double> xyz(3 * n), w(n), txyz(3 * m);
            for (size_t i = 0; i < n; i++) {
                double r = std::sqrt(0.64 + 0.36 * U(g)), ph = 2 * M_PI * U(g), z = U(g) - 0.5;
                xyz[3 * i] = r * std::cos(ph); xyz[3 * i + 1] = r * std::sin(ph); xyz[3 * i + 2] = z;
                w[i] = 1.0 / n;
            }
            for (auto& x : txyz) x = 4 * U(g) - 2;
            CK(fmm_set_sources(c, n, xyz.data(), w.data(), nullptr));
            CK(fmm_set_targets(c, m, txyz.data()));
            double *dB, *dB2, *dt, *dw;
            if (hipMalloc(&dB, 3 * m * 8) || hipMalloc(&dB2, 3 * m * 8) || hipMalloc(&dt, 3 * m * 8) || hipMalloc(&dw, n * 8) ||
                hipMemcpy(dt, txyz.data(), 3 * m * 8, hipMemcpyHostToDevice)) {
                std::printf("HIP allocation failed\n");
                std::_Exit(1);
            }
            for (int it = 0; it < 3; it++) CK(fmm_run_B_dev(c, 1.0, dB));
            CK(fmm_synchronize(c));
            CK(fmm_adjoint_weights_dev(c, 1.0, dB, dw));
            CK(fmm_synchronize(c));
            CK(fmm_direct_B_noalloc_dev(c, 1.0, 4096, dt, dB2));
            CK(fmm_synchronize(c));
            std::vector<double> B(3 * m), B2(3 * 4096);
            if (hipMemcpy(B.data(), dB, 3 * m * 8, hipMemcpyDeviceToHost) ||
                hipMemcpy(B2.data(), dB2, 3 * 4096 * 8, hipMemcpyDeviceToHost)) {
                std::printf("HIP copy failed\n");
                std::_Exit(1);
            }
            double num = 0, den = 0;
            for (size_t k = 0; k < 3 * 4096; k++) { num += (B[k] - B2[k]) * (B[k] - B2[k]); den += B2[k] * B2[k]; }
            fmm_stats st;
            CK(fmm_get_stats(c, &st));
            std::printf("p=%d precision %d: rel-L2 vs direct %.3e, %d source nodes, %.2f ms\n", order, prec,
                        std::sqrt(num / den), st.source_nodes, st.t_total);
            (void)hipFree(dB); (void)hipFree(dB2); (void)hipFree(dt); (void)hipFree(dw);
            fmm_destroy(c);
        }
    }
    std::printf("asan driver done\n");
    std::_Exit(0);   // skips the HIP runtime's exit finalizers, which trip a check in the sanitizer's device allocator hook
}

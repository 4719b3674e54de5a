// fast_multipole_solver.hpp -- source-compatible C++ front end of the MI355X FMM.
//
// Mirrors the reference class FastMultipoleSolver (FMM/FMMGPU/fast_multipole_solver.hpp:13-98,
// base MultipoleSolver multipole_solver.hpp:12-109): same constructor arguments, same method
// names (typo "calclMultipoleExpansions" included) and order of calls, same result type of
// calcB (pairs of point and B in target-leaf order).  Every method forwards to the C ABI of
// include/fmm_capi.h; nothing here computes.  Header only: link libbachelors_fmm.so.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../fmm_capi.h"

namespace bachelors_amd {

struct Vector3 {   // vector3.cuh:8-60 (the members a caller uses)
    double x = 0, y = 0, z = 0;
    Vector3() = default;
    Vector3(double x_, double y_, double z_ = 0) : x(x_), y(y_), z(z_) {}
};

struct Quadrature : Vector3 {   // quadrature.hpp:4-10: weight = volume * basis weight
    double weight = 0;
    Quadrature() = default;
    Quadrature(const Vector3& c, double volume, double w) : Vector3(c), weight(volume * w) {}
};

enum class M2MAlg { NoTranslation = 0, ComplexTranslation, RealTranslation, Layers, Matrices };   // enums :3-10
enum class Device { CPU = 0, GPU, Adaptive };
enum class M2LAlg { ComplexTranslation = 0, Matrices };
enum class Problem { BioSavartLaplace = 0, BEM };

class FastMultipoleSolver {
public:
    bool log = false;
    const int harmonicOrder;
    const int harmonicLength;
    const size_t quadratureOctreeLeafCapacity;
    const size_t calculationPointOctreeLeafCapacity;

    // fast_multipole_solver.hpp:25-30; order / precision are additive runtime options
    FastMultipoleSolver(std::vector<Quadrature>& quadratures, std::vector<Vector3>& points,
                        Problem problem = Problem::BioSavartLaplace, size_t quadratureOctreeLeafCapacity_ = 1000,
                        size_t calculationPointOctreeLeafCapacity_ = 100, int order = 10,
                        int precision = FMM_PREC_FP64)
        : harmonicOrder(order), harmonicLength((order + 1) * (order + 1)),
          quadratureOctreeLeafCapacity(quadratureOctreeLeafCapacity_),
          calculationPointOctreeLeafCapacity(calculationPointOctreeLeafCapacity_), _points(points) {
        if (problem != Problem::BioSavartLaplace) throw std::runtime_error("only BioSavartLaplace is implemented");
        fmm_config cfg;
        fmm_config_init(&cfg);
        cfg.order = order;
        cfg.source_leaf_capacity = (int)quadratureOctreeLeafCapacity_;
        cfg.target_leaf_capacity = (int)calculationPointOctreeLeafCapacity_;
        cfg.precision = precision;
        check(fmm_create(&_ctx, &cfg), "fmm_create");
        std::vector<double> xyz(3 * quadratures.size()), w(quadratures.size());
        for (size_t i = 0; i < quadratures.size(); i++) {
            xyz[3 * i] = quadratures[i].x;
            xyz[3 * i + 1] = quadratures[i].y;
            xyz[3 * i + 2] = quadratures[i].z;
            w[i] = quadratures[i].weight;
        }
        std::vector<double> t(3 * points.size());
        for (size_t i = 0; i < points.size(); i++) {
            t[3 * i] = points[i].x;
            t[3 * i + 1] = points[i].y;
            t[3 * i + 2] = points[i].z;
        }
        check(fmm_set_sources(_ctx, quadratures.size(), xyz.data(), w.data(), nullptr), "fmm_set_sources");
        check(fmm_set_targets(_ctx, points.size(), t.data()), "fmm_set_targets");
        check(fmm_build_trees(_ctx), "fmm_build_trees");   // initTrees (fast_multipole_solver.cpp:43-60)
    }
    FastMultipoleSolver(const FastMultipoleSolver&) = delete;
    FastMultipoleSolver& operator=(const FastMultipoleSolver&) = delete;
    ~FastMultipoleSolver() { fmm_destroy(_ctx); }

    void calcMultipoleExpansionsAtLeaves() {   // multipole_solver.cpp:56-95
        check(fmm_leaf_expansions(_ctx), "calcMultipoleExpansionsAtLeaves");
        _leaves = true;
    }
    // every M2M algorithm and device of the reference yields the same expansions; all run on MFMA here
    void calclMultipoleExpansions(M2MAlg = M2MAlg::Matrices, Device = Device::CPU) {   // :102-139
        if (!_leaves) calcMultipoleExpansionsAtLeaves();
        check(fmm_upward(_ctx), "calclMultipoleExpansions");
    }
    void calcLocalMultipoleExpansions(M2LAlg = M2LAlg::ComplexTranslation, Device = Device::CPU) {   // :62-102
        check(fmm_downward(_ctx), "calcLocalMultipoleExpansions");
    }
    // fast_multipole_solver.cpp:214-231: results in target-leaf order
    std::vector<std::pair<Vector3, Vector3>> calcB(double current) {
        std::vector<double> B(3 * _points.size());
        check(fmm_evaluate_B(_ctx, current, B.data()), "calcB");
        fmm_stats st;
        check(fmm_get_stats(_ctx, &st), "fmm_get_stats");
        std::vector<int64_t> order(st.target_slots);
        check(fmm_leaf_order(_ctx, order.data()), "fmm_leaf_order");
        std::vector<std::pair<Vector3, Vector3>> res;
        res.reserve(order.size());
        for (int64_t i : order) res.emplace_back(_points[i], Vector3(B[3 * i], B[3 * i + 1], B[3 * i + 2]));
        return res;
    }
    fmm_ctx* context() { return _ctx; }

private:
    std::vector<Vector3>& _points;   // non-owning, as the reference (fast_multipole_solver.hpp:16)
    fmm_ctx* _ctx = nullptr;
    bool _leaves = false;
    void check(int rc, const char* what) {
        if (rc != FMM_OK)
            throw std::runtime_error(std::string(what) + ": " + (_ctx ? fmm_last_error(_ctx) : "create failed"));
    }
};

}  // namespace bachelors_amd

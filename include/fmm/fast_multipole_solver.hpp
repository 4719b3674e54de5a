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

struct BEMQuadrature : Quadrature {   // bem_quadrature.hpp: a quadrature with a field B and a normal
    Vector3 B, normal;
    BEMQuadrature() = default;
    BEMQuadrature(const Vector3& c, double area, double w, const Vector3& B_, const Vector3& normal_)
        : Quadrature(c, area, w), B(B_), normal(normal_) {}
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

    // fast_multipole_solver.hpp:25-30; order, precision, tree frame and near-field rule are additive
    // runtime options whose defaults keep the reference's trees and output order (the reference
    // frame) with the corrected numerics (direct near field, analytic curl)
    FastMultipoleSolver(std::vector<Quadrature>& quadratures, std::vector<Vector3>& points,
                        Problem problem = Problem::BioSavartLaplace, size_t quadratureOctreeLeafCapacity_ = 1000,
                        size_t calculationPointOctreeLeafCapacity_ = 100, int order = 10,
                        int precision = FMM_PREC_FP64, int frame = FMM_FRAME_REFERENCE,
                        int near_field = FMM_NEAR_P2P)
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
        cfg.frame = frame;
        cfg.near_field = near_field;
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
    // fast_multipole_solver.hpp:32-37: Problem::BEM with BEMQuadrature sources; their tree is rooted
    // at Box((0,0,0), (1.1, 1.1, 1.6)) (multipole_solver.cpp:668-669)
    FastMultipoleSolver(std::vector<BEMQuadrature>& quadratures, std::vector<Vector3>& points,
                        Problem problem = Problem::BEM, size_t quadratureOctreeLeafCapacity_ = 1000,
                        size_t calculationPointOctreeLeafCapacity_ = 100, int order = 10,
                        int precision = FMM_PREC_FP64, int frame = FMM_FRAME_REFERENCE,
                        int near_field = FMM_NEAR_P2P)
        : harmonicOrder(order), harmonicLength((order + 1) * (order + 1)),
          quadratureOctreeLeafCapacity(quadratureOctreeLeafCapacity_),
          calculationPointOctreeLeafCapacity(calculationPointOctreeLeafCapacity_), _points(points) {
        if (problem != Problem::BEM) throw std::runtime_error("BEMQuadrature sources are for Problem::BEM");
        fmm_config cfg;
        fmm_config_init(&cfg);
        cfg.order = order;
        cfg.source_leaf_capacity = (int)quadratureOctreeLeafCapacity_;
        cfg.target_leaf_capacity = (int)calculationPointOctreeLeafCapacity_;
        cfg.precision = precision;
        cfg.frame = frame;
        cfg.near_field = near_field;
        const double half[3] = {1.1, 1.1, 1.6};
        for (int a = 0; a < 3; a++) { cfg.source_root_center[a] = 0.0; cfg.source_root_half[a] = half[a]; }
        check(fmm_create(&_ctx, &cfg), "fmm_create");
        const size_t n = quadratures.size();
        std::vector<double> xyz(3 * n), w(n), B(3 * n), nrm(3 * n);
        for (size_t i = 0; i < n; i++) {
            const BEMQuadrature& q = quadratures[i];
            xyz[3 * i] = q.x; xyz[3 * i + 1] = q.y; xyz[3 * i + 2] = q.z;
            w[i] = q.weight;
            B[3 * i] = q.B.x; B[3 * i + 1] = q.B.y; B[3 * i + 2] = q.B.z;
            nrm[3 * i] = q.normal.x; nrm[3 * i + 1] = q.normal.y; nrm[3 * i + 2] = q.normal.z;
        }
        std::vector<double> t(3 * points.size());
        for (size_t i = 0; i < points.size(); i++) {
            t[3 * i] = points[i].x; t[3 * i + 1] = points[i].y; t[3 * i + 2] = points[i].z;
        }
        check(fmm_set_bem_sources(_ctx, n, xyz.data(), w.data(), B.data(), nrm.data()), "fmm_set_bem_sources");
        check(fmm_set_targets(_ctx, points.size(), t.data()), "fmm_set_targets");
        check(fmm_build_trees(_ctx), "fmm_build_trees");
    }
    FastMultipoleSolver(const FastMultipoleSolver&) = delete;
    FastMultipoleSolver& operator=(const FastMultipoleSolver&) = delete;
    ~FastMultipoleSolver() { fmm_destroy(_ctx); }

    void calcMultipoleExpansionsAtLeaves() {   // multipole_solver.cpp:56-95
        check(fmm_leaf_expansions(_ctx), "calcMultipoleExpansionsAtLeaves");
        _leaves = true;
    }
    // multipole_solver.cpp:97-100: every node of the source tree, empty children included
    size_t getQuadratureOctreeNodeCount() {
        int nodes = 0, slots = 0;
        check(fmm_tree_size(_ctx, 0, &nodes, &slots), "getQuadratureOctreeNodeCount");
        return (size_t)nodes;
    }
    // multipole_solver.cpp:102-139: every algorithm runs on the GPU (Matrices: MFMA GEMMs;
    // Complex/Real/Layers: the per-node translation formula; NoTranslation: P2M at every node)
    // and gives the same expansions; the device argument is kept for source compatibility
    void calclMultipoleExpansions(M2MAlg alg = M2MAlg::Matrices, Device = Device::CPU) {
        if (!_leaves) calcMultipoleExpansionsAtLeaves();
        check(fmm_upward_with(_ctx, (int)alg), "calclMultipoleExpansions");
    }
    void calcLocalMultipoleExpansions(M2LAlg = M2LAlg::ComplexTranslation, Device = Device::CPU) {   // :62-102
        check(fmm_downward(_ctx), "calcLocalMultipoleExpansions");
    }
    // fast_multipole_solver.cpp:214-231: results in target-leaf order
    std::vector<std::pair<Vector3, Vector3>> calcB(double current) {
        std::vector<double> B(3 * _points.size());
        check(fmm_evaluate_B(_ctx, current, B.data()), "calcB");
        return leaf_ordered(B);
    }
    // fast_multipole_solver.cpp:195-212: A = current / (4 pi) sum w J / |x - y|, target-leaf order
    std::vector<std::pair<Vector3, Vector3>> calcA(double current) {
        std::vector<double> A(3 * _points.size());
        check(fmm_evaluate_A(_ctx, current, A.data()), "calcA");
        return leaf_ordered(A);
    }
    // fast_multipole_solver.cpp:233-250: BEM sources, target-leaf order (current does not enter)
    std::vector<std::pair<Vector3, Vector3>> calcBEM(double /*current*/) {
        std::vector<double> V(3 * _points.size());
        check(fmm_evaluate_bem(_ctx, V.data()), "calcBEM");
        return leaf_ordered(V);
    }
    // multipole_solver.cpp:686-700: the treecode at one point (after calclMultipoleExpansions)
    Vector3 calcB(double current, const Vector3& point) {
        double x[3] = {point.x, point.y, point.z}, B[3];
        check(fmm_treecode(_ctx, current, 1, x, B, nullptr), "calcB(current, point)");
        return Vector3(B[0], B[1], B[2]);
    }
    Vector3 calcA(double current, const Vector3& point) {
        double x[3] = {point.x, point.y, point.z}, A[3];
        check(fmm_treecode(_ctx, current, 1, x, nullptr, A), "calcA(current, point)");
        return Vector3(A[0], A[1], A[2]);
    }
    fmm_ctx* context() { return _ctx; }

private:
    std::vector<std::pair<Vector3, Vector3>> leaf_ordered(const std::vector<double>& v) {
        int nodes = 0, slots = 0;
        check(fmm_tree_size(_ctx, 1, &nodes, &slots), "fmm_tree_size");
        std::vector<int64_t> order(slots);
        check(fmm_leaf_order(_ctx, order.data()), "fmm_leaf_order");
        std::vector<std::pair<Vector3, Vector3>> res;
        res.reserve(order.size());
        for (int64_t i : order) res.emplace_back(_points[i], Vector3(v[3 * i], v[3 * i + 1], v[3 * i + 2]));
        return res;
    }
    std::vector<Vector3>& _points;   // non-owning, as the reference (fast_multipole_solver.hpp:16)
    fmm_ctx* _ctx = nullptr;
    bool _leaves = false;
    void check(int rc, const char* what) {
        if (rc != FMM_OK)
            throw std::runtime_error(std::string(what) + ": " + (_ctx ? fmm_last_error(_ctx) : "create failed"));
    }
};

}  // namespace bachelors_amd

// Drop-in usage exactly as the reference harness drives FastMultipoleSolver
// (FMM/FMMGPU/main.cpp:978-985), against the MI355X library.
#include <cmath>
#include <cstdio>
#include <random>

#include "fmm/fast_multipole_solver.hpp"

using namespace bachelors_amd;

int main() {
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<Quadrature> quads;
    for (int i = 0; i < 4096; i++) {
        double r = std::sqrt(0.64 + 0.36 * U(rng)), ph = 2 * M_PI * U(rng), z = U(rng) - 0.5;
        quads.emplace_back(Vector3(r * std::cos(ph), r * std::sin(ph), z), 1.0, 1.0 / 4096);
    }
    std::vector<Vector3> points;
    for (int i = 0; i < 512; i++) points.emplace_back(4 * U(rng) - 2, 4 * U(rng) - 2, 4 * U(rng) - 2);
    FastMultipoleSolver s(quads, points, Problem::BioSavartLaplace, 32, 32, 12);
    s.calcMultipoleExpansionsAtLeaves();
    s.calclMultipoleExpansions(M2MAlg::Matrices, Device::GPU);
    s.calcLocalMultipoleExpansions(M2LAlg::ComplexTranslation, Device::CPU);
    auto res = s.calcB(5.0);
    // compare with the direct Biot-Savart sum at the result points
    double num = 0, den = 0;
    for (auto& [x, B] : res) {
        double d[3] = {0, 0, 0};
        for (auto& q : quads) {
            double len = std::sqrt(q.x * q.x + q.y * q.y);
            double J[3] = {-q.y / len, q.x / len, 0};
            double r[3] = {x.x - q.x, x.y - q.y, x.z - q.z};
            double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], r3 = r2 * std::sqrt(r2);
            d[0] += q.weight * (J[1] * r[2] - J[2] * r[1]) / r3;
            d[1] += q.weight * (J[2] * r[0] - J[0] * r[2]) / r3;
            d[2] += q.weight * (J[0] * r[1] - J[1] * r[0]) / r3;
        }
        double sc = 5.0 * 1.2566370614e-6 / (4 * 3.14159265359);
        double e[3] = {B.x - sc * d[0], B.y - sc * d[1], B.z - sc * d[2]};
        num += e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        den += sc * sc * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    }
    double rel = std::sqrt(num / den);
    std::printf("drop-in: %zu results, rel-L2 vs direct %.3e\n", res.size(), rel);
    // calcA (same leaf order) and the treecode at one far point (multipole_solver.cpp:686-700)
    auto resA = s.calcA(5.0);
    bool same_order = resA.size() == res.size();
    for (size_t k = 0; same_order && k < res.size(); k++)
        same_order = resA[k].first.x == res[k].first.x && resA[k].first.y == res[k].first.y;
    Vector3 far(10, 10, 10);
    Vector3 Bt = s.calcB(5.0, far);
    double d[3] = {0, 0, 0};
    for (auto& q : quads) {
        double len = std::sqrt(q.x * q.x + q.y * q.y);
        double J[3] = {-q.y / len, q.x / len, 0};
        double r[3] = {far.x - q.x, far.y - q.y, far.z - q.z};
        double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], r3 = r2 * std::sqrt(r2);
        d[0] += q.weight * (J[1] * r[2] - J[2] * r[1]) / r3;
        d[1] += q.weight * (J[2] * r[0] - J[0] * r[2]) / r3;
        d[2] += q.weight * (J[0] * r[1] - J[1] * r[0]) / r3;
    }
    double sc = 5.0 * 1.2566370614e-6 / (4 * 3.14159265359);
    double en = std::hypot(std::hypot(Bt.x - sc * d[0], Bt.y - sc * d[1]), Bt.z - sc * d[2]);
    double dn = sc * std::hypot(std::hypot(d[0], d[1]), d[2]);
    std::printf("drop-in: calcA %zu results, treecode B at (10,10,10) rel err %.3e, %zu source nodes\n", resA.size(),
                en / dn, s.getQuadratureOctreeNodeCount());
    return (res.size() == points.size() && rel < 1e-5 && same_order && en < 1e-6 * dn) ? 0 : 1;
}

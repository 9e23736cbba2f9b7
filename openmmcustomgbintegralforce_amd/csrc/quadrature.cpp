// Quadrature tables of the volume integral: Gauss-Legendre radial nodes on [r0, r1]
// times a Lebedev-Laikov angular rule.  Replaces CharmmQuadrature::GaussLegendre /
// ::Lebedev (openmmapi/include/openmm/internal/CharmmQuadrature.h:14-45) and the table
// assembly of GBSWIntegral.cpp:41-56.
#include "quadrature.h"

#include <cmath>
#include <stdexcept>

namespace gbi {

namespace {

struct Orbit {
    int kind;    // 1: axes, 2: face diagonals, 3: cube diagonals, 4: (a,a,b), 5: (a,b,0)
    double a;    // orbit parameter (kinds 4 and 5)
    double w;    // weight of each point (rule weights sum to 1)
};

void emit(std::vector<std::array<double, 4>>& out, double x, double y, double z, double w) {
    out.push_back({x, y, z, w});
}

// All sign/permutation images of one octahedral orbit.
void expand_orbit(const Orbit& o, std::vector<std::array<double, 4>>& out) {
    const double s[2] = {-1.0, 1.0};
    switch (o.kind) {
    case 1:
        for (double a : s) {
            emit(out, a, 0, 0, o.w);
            emit(out, 0, a, 0, o.w);
            emit(out, 0, 0, a, o.w);
        }
        break;
    case 2: {
        const double c = std::sqrt(0.5);
        for (double a : s)
            for (double b : s) {
                emit(out, 0, a * c, b * c, o.w);
                emit(out, a * c, 0, b * c, o.w);
                emit(out, a * c, b * c, 0, o.w);
            }
        break;
    }
    case 3: {
        const double c = std::sqrt(1.0 / 3.0);
        for (double a : s)
            for (double b : s)
                for (double d : s) emit(out, a * c, b * c, d * c, o.w);
        break;
    }
    case 4: {
        const double b = std::sqrt(1.0 - 2.0 * o.a * o.a);
        for (double u : s)
            for (double v : s)
                for (double t : s) {
                    emit(out, u * o.a, v * o.a, t * b, o.w);
                    emit(out, u * o.a, v * b, t * o.a, o.w);
                    emit(out, u * b, v * o.a, t * o.a, o.w);
                }
        break;
    }
    case 5: {
        const double b = std::sqrt(1.0 - o.a * o.a);
        for (double u : s)
            for (double v : s) {
                emit(out, u * o.a, v * b, 0, o.w);
                emit(out, u * b, v * o.a, 0, o.w);
                emit(out, u * o.a, 0, v * b, o.w);
                emit(out, u * b, 0, v * o.a, o.w);
                emit(out, 0, u * o.a, v * b, o.w);
                emit(out, 0, u * b, v * o.a, o.w);
            }
        break;
    }
    default:
        throw std::invalid_argument("unknown Lebedev orbit");
    }
}

}  // namespace

std::vector<std::array<double, 2>> gauss_legendre(double x0, double x1, int n) {
    if (n <= 0) throw std::invalid_argument("Gauss-Legendre order must be positive");
    std::vector<std::array<double, 2>> out(n);
    const double pi = 3.14159265358979323846;
    const double mid = 0.5 * (x0 + x1), half = 0.5 * (x1 - x0);
    auto pn = [n](double z, double& dp) {
        double a = 1.0, b = z;  // P_{k-1}, P_k
        for (int k = 2; k <= n; ++k) {
            const double c = ((2 * k - 1) * z * b - (k - 1) * a) / k;
            a = b;
            b = c;
        }
        if (n == 1) a = 1.0;
        dp = n * (z * b - a) / (z * z - 1.0);
        return b;
    };
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            const double step = pn(z, dp) / dp;
            z -= step;
            if (std::fabs(step) < 1e-15) break;
        }
        pn(z, dp);
        const double w = 2.0 * half / ((1.0 - z * z) * dp * dp);
        out[i] = {mid - half * z, w};
        out[n - 1 - i] = {mid + half * z, w};
    }
    return out;
}

std::vector<std::array<double, 4>> lebedev(int npts) {
    std::vector<Orbit> orbits;
    switch (npts) {
    case 6: orbits = {{1, 0.0, 1.0 / 6.0}}; break;
    case 14: orbits = {{1, 0.0, 1.0 / 15.0}, {3, 0.0, 3.0 / 40.0}}; break;
    case 26: orbits = {{1, 0.0, 1.0 / 21.0}, {2, 0.0, 4.0 / 105.0}, {3, 0.0, 9.0 / 280.0}}; break;
    case 38:
        orbits = {{1, 0.0, 1.0 / 105.0}, {3, 0.0, 9.0 / 280.0}, {5, 0.4597008433809831, 1.0 / 35.0}};
        break;
    case 50:
        orbits = {{1, 0.0, 4.0 / 315.0},
                  {2, 0.0, 64.0 / 2835.0},
                  {3, 0.0, 27.0 / 1280.0},
                  {4, 0.3015113445777636, 14641.0 / 725760.0}};
        break;
    default:
        throw std::invalid_argument("supported Lebedev rules: 6, 14, 26, 38, 50 points");
    }
    std::vector<std::array<double, 4>> out;
    out.reserve(npts);
    for (const auto& o : orbits) expand_orbit(o, out);
    if ((int)out.size() != npts) throw std::logic_error("Lebedev rule size mismatch");
    return out;
}

std::vector<std::array<double, 5>> quadrature_table(int n_radial, double r0, double r1, int n_lebedev) {
    const auto radial = gauss_legendre(r0, r1, n_radial);
    const auto ang = lebedev(n_lebedev);
    std::vector<std::array<double, 5>> t;
    t.reserve((size_t)n_radial * n_lebedev);
    for (const auto& rw : radial)
        for (const auto& u : ang) t.push_back({rw[0] * u[0], rw[0] * u[1], rw[0] * u[2], rw[0], rw[1] * u[3]});
    return t;
}

}  // namespace gbi

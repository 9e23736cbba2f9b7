// Quadrature tables (host).  Replaces CharmmQuadrature (CharmmQuadrature.h:14-45).
#pragma once

#include <array>
#include <vector>

namespace gbi {

// n Gauss-Legendre nodes (ascending) and weights on [x0, x1].
std::vector<std::array<double, 2>> gauss_legendre(double x0, double x1, int n);
// Lebedev-Laikov rule of npts (6, 14, 26, 38, 50) unit directions; weights sum to 1.
std::vector<std::array<double, 4>> lebedev(int npts);
// Q x (x, y, z, r, w) with index radial * n_lebedev + angular (GBSWIntegral.cpp:41-56).
std::vector<std::array<double, 5>> quadrature_table(int n_radial, double r0, double r1, int n_lebedev);

}  // namespace gbi

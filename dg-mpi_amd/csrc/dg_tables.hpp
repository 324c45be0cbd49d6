// Reference-element tables of the tensor-product hex bases.
#pragma once

#include <vector>

#include "dg_config.hpp"
#include "dg_kernels.hpp"

namespace dg {

// n-point Gauss-Legendre rule on [-1,1], ascending (the reference's
// tabulated rules, src/numerics/nodes.cpp:97-500 and
// src/numerics/quadrature/segment.cpp:7-411, to the nearest double).
void gl_rule(int n, double* x, double* w);

// 1D factors phi_i(x), phi_i'(x) of LagrangeHex (equidistant nodes,
// src/numerics/basis/tools.cpp:9-25,115-149) or LegendreHex (unnormalised
// P_n, src/numerics/basis/tools.cpp:298-366).
void basis_1d(Basis b, int p, double x, double* val, double* grad);

// Kernel tables for order p (p <= PMAX).
KTables make_tables(const CaseConfig& c);

// Structured-box geometry and physics constants.
KGeom make_geometry(const CaseConfig& c);

KIC make_ic(const CaseConfig& c);

// Node (i, j, k), i in [0, nelem[0]] etc., of the periodic box, displaced
// by Mesh.warp:
//   x = lo_x + i h_x + warp h_x sin(2 pi eta) sin(2 pi zeta)   (y, z cyclic)
// with xi = (i mod nx)/nx etc.  Each component depends on the other two
// indices only, so opposite box faces stay translates of each other.
void box_node(const CaseConfig& c, long i, long j, long k, double* x);

// Node coordinates of every element, [E][8][3], node n = 4k + 2j + i
// (the reference's hex node order), element e = ix + nx (iy + ny iz).
std::vector<double> element_nodes(const CaseConfig& c);

}  // namespace dg

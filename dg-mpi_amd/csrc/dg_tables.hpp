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

// Face records and normal-evaluation points of the general path (the
// layout KGeom::frec / KGeom::nev describe).
struct FaceTables {
    std::vector<unsigned> frec;   // [E][6][2]
    std::vector<double> nev;      // [48][(p+2)^2][2]
};

// Records of the periodic box (structured neighbours, aligned frames).
FaceTables box_face_tables(const CaseConfig& c);

// Records of a mesh given by the reference's IFaceData 6-tuples (elemL,
// faceL, orientL, elemR, faceR, orientR), src/mesh/mesh.cpp:140-160.  Every
// element face must be an interior face (periodic, no boundary faces).
FaceTables mesh_face_tables(const CaseConfig& c, long E, long F, const int* iface6);

// The same for quadrilateral meshes: frec [E][4][2], slot d*2 + side
// (d = 0: faces 3, 1; d = 1: faces 0, 2): neighbour element, then the
// neighbour's face id (bits 0-1) | the face points run in opposite senses of
// the two tangential coordinates (bit 2) | this element is the left one
// (bit 3).  No normal table: the faces are straight.
FaceTables mesh_face_tables_2d(const CaseConfig& c, long E, long F, const int* iface6);

}  // namespace dg

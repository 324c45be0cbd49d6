/*
 * dg_oracle.c -- CPU oracle (TEST INFRASTRUCTURE) for the DG residual + RK
 * path of IhmeGroup/DG-MPI: 3D Euler, periodic hexahedral boxes, FP64.
 *
 * A restatement of the reference algorithm in plain C99 + OpenMP.  It keeps
 * the reference's dense formulation (per-element GEMMs against tabulated
 * basis matrices, gathered face states with orientation index maps, a dense
 * LU-inverted mass matrix per element, the BLAS-1 RK4 sequence) so that it
 * checks the sum-factorised HIP kernels against an independent route.
 *
 * Only tests/, the smoke check and the benchmark's CPU-baseline leg may use
 * this file.  Each function cites the reference source it restates (paths
 * are relative to the reference repository root).
 *
 * Pins (see tests/test_oracle_*.py):
 *   - GL nodes/weights: the reference tables (src/numerics/nodes.cpp,
 *     src/numerics/quadrature/segment.cpp) transcribed as a fixture, plus
 *     weight sums (test/numerics/quadrature/test_quadrature.cpp:19-122).
 *   - Legendre basis: test/numerics/basis/test_basis.cpp:308-686 vectors.
 *   - Lagrange basis: nodal property and p=1 gradients (test_basis.cpp).
 *   - physics: closed forms of test/test_physics.cpp:18-110,344-443 and
 *     the flux consistency/conservation properties at :479-576.
 *   - face pairing: physical coincidence of paired face points.
 *   - mass matrix: M*Minv = I, Legendre closed form 2/(2i+1).
 *   - residual: free-stream preservation and discrete conservation.
 *   - end to end: examples/euler/3D/isentropic_vortex/regression_data.txt.
 * Parity unpinned: the operation order inside KokkosKernels' batched GEMM /
 * LU (the library is absent from the reference mount); restated here as
 * the textbook unblocked algorithms.
 */
#include "dg_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NS ORC_NS

static void die(const char* msg) {
    fprintf(stderr, "dg_oracle: %s\n", msg);
    abort();
}

static void* xcalloc(size_t n, size_t sz) {
    void* p = calloc(n ? n : 1, sz);
    if (!p) die("out of memory");
    return p;
}

/* ======================================================================
 * Quadrature
 * ==================================================================== */

/* src/numerics/quadrature/tools.cpp:20-24 */
int orc_nq1d_from_qorder(int order) {
    int nq_1d = 0;
    if ((order + 2) % 2 == 0) nq_1d = (order + 2) / 2;
    if ((order + 1) % 2 == 0) nq_1d = (order + 1) / 2;
    return nq_1d;
}

/* src/numerics/quadrature/tools.cpp:7-17 (hexahedron; gorder fixed to 1) */
int orc_hex_qorder(int order) { return 2 * (order + 1) + 1; }

/* src/numerics/quadrature/quadrilateral.cpp:100-105 with
 * src/numerics/quadrature/segment.cpp:423-425 */
int orc_quad_face_qorder(int order) { return (2 * order + 1) + 2; }

/* Gauss-Legendre nodes/weights on [-1,1], ascending.  The reference stores
 * 32-digit decimal tables (src/numerics/nodes.cpp:97-500,
 * src/numerics/quadrature/segment.cpp:7-411) selected by case(2k,2k+1) ->
 * k+1 points; here they are computed by Newton's method on P_n in binary128
 * and rounded, which reproduces the correctly rounded table literals.  The
 * tables are symmetric with an exact 0 in the middle; so is this. */
void orc_gl_rule(int n, double* pts, double* wts) {
    if (n < 1) die("gl_rule: n < 1");
    if (n == 1) { pts[0] = 0.0; wts[0] = 2.0; return; }
    for (int i = 0; i < n / 2; ++i) {
        __float128 z = cos(M_PI * (i + 0.75) / (n + 0.5));
        __float128 dp = 0;
        for (int it = 0; it < 100; ++it) {
            __float128 p0 = 1, p1 = z;
            for (int k = 2; k <= n; ++k) {
                __float128 p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1; p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1);
            __float128 dz = p1 / dp;
            z -= dz;
            if ((dz < 0 ? -dz : dz) < 1e-33) {
                /* recompute derivative at the converged root */
                p0 = 1; p1 = z;
                for (int k = 2; k <= n; ++k) {
                    __float128 p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                    p0 = p1; p1 = p2;
                }
                dp = n * (z * p1 - p0) / (z * z - 1);
                break;
            }
        }
        __float128 w = 2 / ((1 - z * z) * dp * dp);
        pts[n - 1 - i] = (double)z;
        pts[i] = -(double)z;
        wts[n - 1 - i] = (double)w;
        wts[i] = (double)w;
    }
    if (n % 2 == 1) {
        int m = n / 2;
        pts[m] = 0.0;
        /* middle weight: P'_n(0) via recurrence in binary128 */
        __float128 p0 = 1, p1 = 0, z = 0;
        for (int k = 2; k <= n; ++k) {
            __float128 p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
            p0 = p1; p1 = p2;
        }
        __float128 dp = n * (z * p1 - p0) / (z * z - 1);
        wts[m] = (double)(2 / ((1 - z * z) * dp * dp));
    }
}

/* Tensor rules: src/numerics/quadrature/hexahedron.cpp:51-61 (q = k m^2 +
 * j m + i, x fastest) and quadrilateral.cpp:90-97 (q = j m + i). */
static void hex_rule(int nq1d, double* pts /* [nq][3] */, double* wts) {
    double x[32], w[32];
    orc_gl_rule(nq1d, x, w);
    for (int k = 0; k < nq1d; ++k)
        for (int j = 0; j < nq1d; ++j)
            for (int i = 0; i < nq1d; ++i) {
                int q = k * nq1d * nq1d + j * nq1d + i;
                pts[3 * q + 0] = x[i];
                pts[3 * q + 1] = x[j];
                pts[3 * q + 2] = x[k];
                wts[q] = w[i] * w[j] * w[k];
            }
}

static void quad_rule(int nq1d, double* pts /* [nq][2] */, double* wts) {
    double x[32], w[32];
    orc_gl_rule(nq1d, x, w);
    for (int j = 0; j < nq1d; ++j)
        for (int i = 0; i < nq1d; ++i) {
            int q = j * nq1d + i;
            pts[2 * q + 0] = x[i];
            pts[2 * q + 1] = x[j];
            wts[q] = w[i] * w[j];
        }
}

/* ======================================================================
 * Basis
 * ==================================================================== */

/* src/numerics/basis/tools.cpp:9-25 */
void orc_equidistant_nodes(double start, double stop, int nnodes, double* x) {
    if (nnodes <= 1) { x[0] = 0.; return; }
    double dx = (stop - start) / ((double)nnodes - 1.);
    for (int i = 0; i < nnodes; i++) x[i] = start + (double)i * dx;
}

/* src/numerics/basis/tools.cpp:115-127 */
static void lagrange_val_1d(double x, const double* xn, int nn, double* phi) {
    for (int i = 0; i < nn; i++) {
        phi[i] = 1.;
        for (int j = 0; j < nn; j++) {
            if (i == j) continue;
            phi[i] *= (x - xn[j]) / (xn[i] - xn[j]);
        }
    }
}

/* src/numerics/basis/tools.cpp:130-149 */
static void lagrange_grad_1d(double x, const double* xn, int nn, double* g) {
    for (int i = 0; i < nn; i++) {
        g[i] = 0.;
        for (int j = 0; j < nn; j++) {
            if (i == j) continue;
            double prod = 1.;
            for (int k = 0; k < nn; k++) {
                if (i == k) continue;
                else if (j == k) prod *= 1. / (xn[i] - xn[k]);
                else prod *= (x - xn[k]) / (xn[i] - xn[k]);
            }
            g[i] += prod;
        }
    }
}

/* src/numerics/basis/tools.cpp:298-331 (unnormalised P_n, p <= 7) */
static void legendre_val_1d(double x, int p, double* phi) {
    if (p > 7) die("Legendre polynomial P > 7 not implemented (reference limit)");
    if (p >= 0) phi[0] = 1.;
    if (p >= 1) phi[1] = x;
    if (p >= 2) phi[2] = 0.5*(3.*x*x - 1.);
    if (p >= 3) phi[3] = 0.5*(5.*x*x*x - 3.*x);
    if (p >= 4) phi[4] = 0.125*(35.*x*x*x*x - 30.*x*x + 3.);
    if (p >= 5) phi[5] = 0.125*(63.*x*x*x*x*x - 70.*x*x*x + 15.*x);
    if (p >= 6) phi[6] = 0.0625*(231.*x*x*x*x*x*x - 315.*x*x*x*x + 105.*x*x -5.);
    if (p == 7) phi[7] = 0.0625*(429.*x*x*x*x*x*x*x - 693.*x*x*x*x*x + 315.*x*x*x - 35.*x);
}

/* src/numerics/basis/tools.cpp:334-366 */
static void legendre_grad_1d(double x, int p, double* g) {
    if (p > 7) die("Legendre polynomial P > 7 not implemented (reference limit)");
    if (p >= 0) g[0] = 0.0;
    if (p >= 1) g[1] = 1.;
    if (p >= 2) g[2] = 3.*x;
    if (p >= 3) g[3] = 0.5*(15.*x*x - 3.);
    if (p >= 4) g[4] = 0.125*(35.*4.*x*x*x - 60.*x);
    if (p >= 5) g[5] = 0.125*(63.*5.*x*x*x*x - 210.*x*x + 15.);
    if (p >= 6) g[6] = 0.0625*(231.*6.*x*x*x*x*x - 315.*4.*x*x*x + 210.*x);
    if (p == 7) g[7] = 0.0625*(429.*7.*x*x*x*x*x*x - 693.*5.*x*x*x*x + 315.*3.*x*x - 35.);
}

/* 1D factors of the hex bases: src/numerics/basis/basis.cpp:93-128
 * (Lagrange on equidistant nodes, basis.cpp:240) and basis.cpp:204-232
 * (Legendre).  order 0: value 1, gradient left at its zero init. */
void orc_basis_1d(int basis, int p, double x, double* val, double* grad) {
    if (p == 0) { val[0] = 1.; grad[0] = 0.; return; }
    if (basis == ORC_LAGRANGE_HEX) {
        double xn[32];
        orc_equidistant_nodes(-1., 1., p + 1, xn);
        lagrange_val_1d(x, xn, p + 1, val);
        lagrange_grad_1d(x, xn, p + 1, grad);
    } else {
        legendre_val_1d(x, p, val);
        legendre_grad_1d(x, p, grad);
    }
}

/* src/numerics/basis/tools.cpp:216-295 (Lagrange) and :430-508 (Legendre):
 * basis index k (p+1)^2 + j (p+1) + i, value vx*vy*vz, gradients
 * (gx*vy*vz, vx*gy*vz, vx*vy*gz). */
void orc_basis_3d(int basis, int p, int npts, const double* pts,
                  double* val, double* grad) {
    int n = p + 1, nb = n * n * n;
    for (int q = 0; q < npts; ++q) {
        double vx[32], vy[32], vz[32], gx[32], gy[32], gz[32];
        orc_basis_1d(basis, p, pts[3 * q + 0], vx, gx);
        orc_basis_1d(basis, p, pts[3 * q + 1], vy, gy);
        orc_basis_1d(basis, p, pts[3 * q + 2], vz, gz);
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    int b = k * n * n + j * n + i;
                    if (val) val[(size_t)q * nb + b] = vx[i] * vy[j] * vz[k];
                    if (grad) {
                        if (p == 0) {
                            grad[((size_t)q * nb + b) * 3 + 0] = 0.;
                            grad[((size_t)q * nb + b) * 3 + 1] = 0.;
                            grad[((size_t)q * nb + b) * 3 + 2] = 0.;
                        } else {
                            grad[((size_t)q * nb + b) * 3 + 0] = gx[i] * vy[j] * vz[k];
                            grad[((size_t)q * nb + b) * 3 + 1] = vx[i] * gy[j] * vz[k];
                            grad[((size_t)q * nb + b) * 3 + 2] = vx[i] * vy[j] * gz[k];
                        }
                    }
                }
    }
}

/* Geometric basis LagrangeHex gorder=1 (src/exec/main.cpp:55-59; node
 * n = 4k + 2j + i): values and gradients at npts points. */
static void gbasis_hex(int npts, const double* pts, double* val, double* grad) {
    orc_basis_3d(ORC_LAGRANGE_HEX, 1, npts, pts, val, grad);
}

/* ======================================================================
 * Hexahedron shape: faces, orientations, normals
 * ==================================================================== */

/* src/numerics/basis/shape.cpp:224-324 */
void orc_points_on_face(int face_id, int orient, int np, const double* face_pts,
                        double* elem_pts) {
    static const double xr[8] = {-1., 1.,-1., 1.,-1., 1.,-1., 1.};
    static const double yr[8] = {-1.,-1., 1., 1.,-1.,-1., 1., 1.};
    static const double zr[8] = {-1.,-1.,-1.,-1., 1., 1., 1., 1.};
    static const int ftab[6][4] = {{0, 2, 3, 1}, {0, 1, 5, 4}, {1, 3, 7, 5},
                                   {3, 2, 6, 7}, {2, 0, 4, 6}, {4, 5, 7, 6}};
    if (face_id < 0 || face_id > 5) die("face id not available for hexahedron");
    int node[4];
    int cycle = orient % 4, flip = orient / 4;
    for (int j = 0; j < 4; j++) node[j] = ftab[face_id][(j + cycle) % 4];
    if (flip >= 1) { int t = node[1]; node[1] = node[3]; node[3] = t; }
    double xe[4], ye[4], ze[4], b[4];
    for (int j = 0; j < 4; j++) { xe[j] = xr[node[j]]; ye[j] = yr[node[j]]; ze[j] = zr[node[j]]; }
    for (int i = 0; i < np; i++) {
        double xb = (face_pts[2 * i + 0] + 1.) / 2.;
        double yb = (face_pts[2 * i + 1] + 1.) / 2.;
        b[0] = (1. - xb) * (1. - yb);
        b[1] = xb * (1. - yb);
        b[2] = xb * yb;
        b[3] = (1. - xb) * yb;
        double ex = 0.0, ey = 0.0, ez = 0.0;
        for (int j = 0; j < 4; j++) { ex += b[j] * xe[j]; ey += b[j] * ye[j]; ez += b[j] * ze[j]; }
        elem_pts[3 * i + 0] = ex; elem_pts[3 * i + 1] = ey; elem_pts[3 * i + 2] = ez;
    }
}

/* src/numerics/basis/shape.cpp:326-428 */
void orc_face_pts_order(int orient, int npts, int* po) {
    int n1D = (int)lround(sqrt((double)npts));
    if (n1D * n1D != npts) die("face point count is not a square");
    int nptsp1 = npts + 1;
    switch (orient) {
    case 0: for (int i = 0; i < npts; i++) po[i] = i; break;
    case 1: { int curr = n1D - 1;
        for (int i = 0; i < npts; i++) { po[i] = curr; curr = (curr + n1D) % nptsp1; } break; }
    case 2: for (int i = 0; i < npts; i++) po[i] = npts - 1 - i; break;
    case 3: { int curr = npts - n1D;
        for (int i = 0; i < npts; i++) { po[i] = curr; curr -= n1D; if (curr < 0) curr += nptsp1; } break; }
    case 4: { int curr = 0;
        for (int i = 0; i < npts; i++) { po[i] = curr; curr += n1D; if (curr >= npts) curr -= npts - 1; } break; }
    case 5: { int curr = n1D - 1, k = 0;
        for (int i = 0; i < n1D; i++) { for (int j = 0; j < n1D; j++) { po[k] = curr; curr -= 1; k += 1; } curr += 2 * n1D; } break; }
    case 6: { int curr = npts - 1;
        for (int i = 0; i < npts; i++) { po[i] = curr; curr -= n1D; if (curr < 0) curr += npts - 1; } break; }
    case 7: { int curr = npts - n1D, k = 0;
        for (int i = 0; i < n1D; i++) { for (int j = 0; j < n1D; j++) { po[k] = curr; curr += 1; k += 1; } curr -= 2 * n1D; } break; }
    default: die("bad face orientation");
    }
}

/* src/numerics/basis/shape.cpp:430-480 */
void orc_local_nodes_on_face(int face_id, int gorder, int* lf) {
    int i0 = 0, d0 = 0, d1 = 0;
    switch (face_id) {
    case 0: i0 = 0; d0 = gorder + 1; d1 = 1; break;
    case 1: i0 = 0; d0 = 1; d1 = (gorder + 1) * (gorder + 1); break;
    case 2: i0 = gorder; d0 = gorder + 1; d1 = (gorder + 1) * (gorder + 1); break;
    case 3: i0 = (gorder + 1) * (gorder + 1) - 1; d0 = -1; d1 = (gorder + 1) * (gorder + 1); break;
    case 4: i0 = gorder * (gorder + 1); d0 = -(gorder + 1); d1 = (gorder + 1) * (gorder + 1); break;
    case 5: i0 = gorder * (gorder + 1) * (gorder + 1); d0 = 1; d1 = gorder + 1; break;
    default: die("bad face id");
    }
    for (int j = 0, k = 0; j < gorder + 1; j++)
        for (int i = 0; i < gorder + 1; i++, k++) lf[k] = i0 + i * d0 + j * d1;
}

/* ======================================================================
 * Dense linear algebra (semantics of src/math/linear_algebra.cpp:75-259;
 * KokkosBatched internals restated as textbook unblocked loops --
 * parity unpinned at the operation-order level).
 * ==================================================================== */

/* C(m x n) = beta*C + alpha * A(m x k) B(k x n); strides given per matrix:
 * A(i,p) = A[i*as0 + p*as1] etc.  beta is 0 or 1. */
static void gemm(int m, int n, int k, double alpha,
                 const double* A, int as0, int as1,
                 const double* B, int bs0, int bs1,
                 double beta, double* C, int cs0, int cs1) {
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            double c = 0;
            for (int p = 0; p < k; ++p) c += A[i * as0 + p * as1] * B[p * bs0 + j * bs1];
            double* cij = &C[i * cs0 + j * cs1];
            *cij = (beta == 0. ? 0. : *cij) + alpha * c;
        }
}

/* det of 3x3: src/math/linear_algebra.cpp:243-259 */
static double det3(const double* m) {
#define M_(i, j) m[(i) * 3 + (j)]
    return M_(0,0)*M_(1,1)*M_(2,2)-M_(0,0)*M_(1,2)*M_(2,1)
         -M_(0,1)*M_(1,0)*M_(2,2)+M_(0,1)*M_(1,2)*M_(2,0)
         +M_(0,2)*M_(1,0)*M_(2,1)-M_(0,2)*M_(1,1)*M_(2,0);
}

/* closed-form 3x3 inverse: src/math/linear_algebra.cpp:150-163 */
static void inv3(const double* m, double* im) {
    double det = det3(m);
#define I_(i, j) im[(i) * 3 + (j)]
    I_(0,0)=(M_(1,1)*M_(2,2)-M_(1,2)*M_(2,1))/det;
    I_(1,0)=-(M_(1,0)*M_(2,2)-M_(1,2)*M_(2,0))/det;
    I_(2,0)=(M_(1,0)*M_(2,1)-M_(1,1)*M_(2,0))/det;
    I_(0,1)=-(M_(0,1)*M_(2,2)-M_(0,2)*M_(2,1))/det;
    I_(1,1)=(M_(0,0)*M_(2,2)-M_(0,2)*M_(2,0))/det;
    I_(2,1)=-(M_(0,0)*M_(2,1)-M_(0,1)*M_(2,0))/det;
    I_(0,2)=(M_(0,1)*M_(1,2)-M_(0,2)*M_(1,1))/det;
    I_(1,2)=-(M_(0,0)*M_(1,2)-M_(0,2)*M_(1,0))/det;
    I_(2,2)=(M_(0,0)*M_(1,1)-M_(0,1)*M_(1,0))/det;
#undef I_
#undef M_
}

/* invA for n > 3: LU without pivoting, then LU X = I
 * (src/math/linear_algebra.cpp:165-174, KokkosBatched SerialLU /
 * SerialSolveLU, unblocked).  n <= 3 closed forms (:139-164). */
static void inv_dense(int n, double* A /* destroyed */, double* X) {
    if (n == 1) { X[0] = 1. / A[0]; return; }
    if (n == 2) {
        double det = A[0] * A[3] - A[1] * A[2];
        X[0] = A[3] / det; X[1] = -A[1] / det; X[2] = -A[2] / det; X[3] = A[0] / det;
        return;
    }
    if (n == 3) { inv3(A, X); return; }
    for (int i = 0; i < n * n; ++i) X[i] = 0.;
    for (int i = 0; i < n; ++i) X[i * n + i] = 1.;
    for (int p = 0; p < n; ++p) {             /* right-looking Doolittle */
        double alpha = A[p * n + p];
        for (int i = p + 1; i < n; ++i) {
            A[i * n + p] /= alpha;
            for (int j = p + 1; j < n; ++j) A[i * n + j] -= A[i * n + p] * A[p * n + j];
        }
    }
    for (int p = 0; p < n; ++p)                /* L (unit) forward, column-oriented */
        for (int i = p + 1; i < n; ++i)
            for (int j = 0; j < n; ++j) X[i * n + j] -= A[i * n + p] * X[p * n + j];
    for (int p = n - 1; p >= 0; --p) {         /* U backward */
        double alpha = A[p * n + p];
        for (int j = 0; j < n; ++j) X[p * n + j] /= alpha;
        for (int i = 0; i < p; ++i)
            for (int j = 0; j < n; ++j) X[i * n + j] -= A[i * n + p] * X[p * n + j];
    }
}

/* ======================================================================
 * Physics
 * ==================================================================== */

/* src/physics/euler/functions.cpp:10-23 (Math::dot<3> order) */
double orc_pressure(double gamma, const double* U) {
    double rKE = (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0];
    return (gamma - 1.0) * (U[4] - 0.5 * rKE);
}

/* src/physics/euler/functions.cpp:25-40 */
double orc_maxwavespeed(double gamma, const double* U) {
    double rho1 = 1. / U[0];
    double sqrtrKE = sqrt(U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * rho1;
    return sqrtrKE + sqrt(gamma * orc_pressure(gamma, U) * rho1);
}

/* src/physics/euler/functions.cpp:64-91; Fdir[d*5 + s] */
void orc_flux_interior(double gamma, const double* U, double* F) {
    const double r1 = 1. / U[0];
    const double ru = U[1], rv = U[2], rw = U[3], rE = U[4];
    const double P = orc_pressure(gamma, U);
    F[0] = ru;  F[1] = ru*ru*r1 + P;  F[2] = ru*rv*r1;  F[3] = ru*rw*r1;  F[4] = (rE + P)*ru*r1;
    F[5] = rv;  F[6] = rv*ru*r1;  F[7] = rv*rv*r1 + P;  F[8] = rv*rw*r1;  F[9] = (rE + P)*rv*r1;
    F[10] = rw; F[11] = rw*ru*r1; F[12] = rw*rv*r1; F[13] = rw*rw*r1 + P; F[14] = (rE + P)*rw*r1;
}

/* src/physics/base/functions.cpp:33-89 */
static void flux_llf(double gamma, const double* UL, const double* UR,
                     const double* N, double* F) {
    double n_mag = 0.0, n_hat[3];
    for (int d = 0; d < 3; d++) n_mag += N[d] * N[d];
    n_mag = sqrt(n_mag);
    for (int d = 0; d < 3; d++) n_hat[d] = N[d] / n_mag;
    double Fq[15], FqL[NS] = {0}, FqR[NS] = {0}, dU[NS];
    orc_flux_interior(gamma, UL, Fq);
    for (int s = 0; s < NS; s++) for (int d = 0; d < 3; d++) FqL[s] += Fq[s + d * NS] * n_hat[d];
    orc_flux_interior(gamma, UR, Fq);
    for (int s = 0; s < NS; s++) for (int d = 0; d < 3; d++) FqR[s] += Fq[s + d * NS] * n_hat[d];
    for (int s = 0; s < NS; s++) dU[s] = UR[s] - UL[s];
    double a = orc_maxwavespeed(gamma, UL);
    double aR = orc_maxwavespeed(gamma, UR);
    if (aR > a) a = aR;
    for (int s = 0; s < NS; s++) F[s] = n_mag * (0.5 * (FqL[s] + FqR[s]) - 0.5 * a * dU[s]);
}

/* src/physics/euler/functions.cpp:298-399 */
static void flux_hllc(double gam, const double* UL, const double* UR,
                      const double* N, double* F) {
    const double djac = sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    const double djac1 = 1. / djac;
    const double n[3] = {N[0] * djac1, N[1] * djac1, N[2] * djac1};
    const double rhoL = UL[0], rhoL1 = 1. / rhoL;
    const double unL = (UL[1] * n[0] + UL[2] * n[1] + UL[3] * n[2]) * rhoL1;
    const double PL = orc_pressure(gam, UL);
    const double aL = sqrt(gam * PL * rhoL1);
    const double rhoR = UR[0], rhoR1 = 1. / rhoR;
    const double unR = (UR[1] * n[0] + UR[2] * n[1] + UR[3] * n[2]) * rhoR1;
    const double PR = orc_pressure(gam, UR);
    const double aR = sqrt(gam * PR * rhoR1);
    const double rAvg = 0.5 * (rhoL + rhoR);
    const double aAvg = 0.5 * (aL + aR);
    const double pStar = fmax(0., 0.5 * (PL + PR - (unR - unL) * rAvg * aAvg));
    double qL = 1.0;
    if (pStar / PL > 1.) qL = sqrt(1. + (gam + 1.) / (2. * gam) * (pStar / PL - 1.));
    double qR = 1.0;
    if (pStar / PR > 1.) qR = sqrt(1. + (gam + 1.) / (2. * gam) * (pStar / PR - 1.));
    const double sL = unL - aL * qL;
    const double sR = unR + aR * qR;
    const double sStar = (PR - PL + rhoL * unL * (sL - unL) - rhoR * unR * (sR - unR)) /
                         (rhoL * (sL - unL) - rhoR * (sR - unR));
    if (sL >= 0.) {
        F[0] = djac * rhoL * unL;
        for (int i = 0; i < 3; i++) F[1 + i] = djac * (UL[1 + i] * unL + PL * n[i]);
        F[4] = djac * ((UL[4] + PL) * unL);
    } else if (sR <= 0.) {
        F[0] = djac * rhoR * unR;
        for (int i = 0; i < 3; i++) F[1 + i] = djac * (UR[1 + i] * unR + PR * n[i]);
        F[4] = djac * (UR[4] + PR) * unR;
    } else if ((sL < 0.) && (sStar >= 0.)) {
        const double c = (sL - unL) / (sL - sStar);
        const double f1 = sStar - unL;
        const double f2 = sStar + PL / (rhoL * (sL - unL));
        F[0] = djac * (rhoL * unL + rhoL * sL * (c - 1.));
        for (int i = 0; i < 3; i++)
            F[1 + i] = djac * (UL[1 + i] * unL + PL * n[i] + sL * (UL[1 + i] * (c - 1.) + rhoL * c * f1 * n[i]));
        F[4] = djac * ((UL[4] + PL) * unL + sL * (UL[4] * (c - 1.) + rhoL * c * f1 * f2));
    } else if ((sStar < 0.) && (sR > 0.)) {
        const double c = (sR - unR) / (sR - sStar);
        const double f1 = sStar - unR;
        const double f2 = sStar + PR / (rhoR * (sR - unR));
        F[0] = djac * (rhoR * unR + rhoR * sR * (c - 1.));
        for (int i = 0; i < 3; i++)
            F[1 + i] = djac * (UR[1 + i] * unR + PR * n[i] + sR * (UR[1 + i] * (c - 1.) + rhoR * c * f1 * n[i]));
        F[4] = djac * ((UR[4] + PR) * unR + sR * (UR[4] * (c - 1.) + rhoR * c * f1 * f2));
    } else {
        /* reference prints and assert(false)s (functions.cpp:389-398) */
        for (int s = 0; s < NS; s++) F[s] = NAN;
    }
}

/* Roe flux: absent from the reference (only a commented call site,
 * src/physics/euler/euler.cpp:846-857).  New: Roe-Pike wave
 * decomposition (Toro, Riemann Solvers, 3rd ed., sec. 11.2-11.3) with a
 * Harten-Hyman-type entropy fix on the two acoustic waves:
 *   delta = max(0, lam~ - lam(UL), lam(UR) - lam~),
 *   |lam~| -> (lam~^2 + delta^2) / (2 delta)  when |lam~| < delta.
 * delta > 0 only across an expansive (transonic rarefaction) wave, so a
 * single shock -- including a stationary one -- stays exact.
 * Same convention as the reference fluxes: N is area weighted, the result
 * is |N| * Fhat(n_hat). */
static double entropy_fix(double lam, double lamL, double lamR) {
    double delta = fmax(0., fmax(lam - lamL, lamR - lam));
    double a = fabs(lam);
    return a < delta ? 0.5 * (lam * lam + delta * delta) / delta : a;
}

static void flux_roe(double gam, const double* UL, const double* UR,
                     const double* N, double* F) {
    const double nmag = sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    const double n[3] = {N[0] / nmag, N[1] / nmag, N[2] / nmag};
    const double rL = UL[0], rR = UR[0];
    const double uL[3] = {UL[1] / rL, UL[2] / rL, UL[3] / rL};
    const double uR[3] = {UR[1] / rR, UR[2] / rR, UR[3] / rR};
    const double pL = orc_pressure(gam, UL), pR = orc_pressure(gam, UR);
    const double HL = (UL[4] + pL) / rL, HR = (UR[4] + pR) / rR;
    const double sL = sqrt(rL), sR = sqrt(rR), isum = 1. / (sL + sR);
    double ut[3];
    for (int d = 0; d < 3; d++) ut[d] = (sL * uL[d] + sR * uR[d]) * isum;
    const double Ht = (sL * HL + sR * HR) * isum;
    const double q2 = ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2];
    const double c2 = (gam - 1.) * (Ht - 0.5 * q2);
    const double ct = sqrt(c2);
    const double rt = sL * sR;
    const double unt = ut[0] * n[0] + ut[1] * n[1] + ut[2] * n[2];
    const double unL = uL[0] * n[0] + uL[1] * n[1] + uL[2] * n[2];
    const double unR = uR[0] * n[0] + uR[1] * n[1] + uR[2] * n[2];
    const double drho = rR - rL, dp = pR - pL, dun = unR - unL;
    double du[3];
    for (int d = 0; d < 3; d++) du[d] = uR[d] - uL[d];
    const double a1 = (dp - rt * ct * dun) / (2. * c2);
    const double a5 = (dp + rt * ct * dun) / (2. * c2);
    const double a2 = drho - dp / c2;
    const double cL = sqrt(gam * pL / rL), cR = sqrt(gam * pR / rR);
    const double l1 = entropy_fix(unt - ct, unL - cL, unR - cR);
    const double l2 = fabs(unt);
    const double l5 = entropy_fix(unt + ct, unL + cL, unR + cR);
    double D[NS];
    /* acoustic waves */
    D[0] = l1 * a1 + l5 * a5;
    for (int d = 0; d < 3; d++) D[1 + d] = l1 * a1 * (ut[d] - ct * n[d]) + l5 * a5 * (ut[d] + ct * n[d]);
    D[4] = l1 * a1 * (Ht - unt * ct) + l5 * a5 * (Ht + unt * ct);
    /* entropy wave */
    D[0] += l2 * a2;
    for (int d = 0; d < 3; d++) D[1 + d] += l2 * a2 * ut[d];
    D[4] += l2 * a2 * 0.5 * q2;
    /* shear waves */
    double udu = 0.;
    for (int d = 0; d < 3; d++) {
        double dut = du[d] - dun * n[d];
        D[1 + d] += l2 * rt * dut;
        udu += ut[d] * dut;
    }
    D[4] += l2 * rt * udu;
    double Fq[15], FL[NS] = {0}, FR[NS] = {0};
    orc_flux_interior(gam, UL, Fq);
    for (int s = 0; s < NS; s++) for (int d = 0; d < 3; d++) FL[s] += Fq[s + d * NS] * n[d];
    orc_flux_interior(gam, UR, Fq);
    for (int s = 0; s < NS; s++) for (int d = 0; d < 3; d++) FR[s] += Fq[s + d * NS] * n[d];
    for (int s = 0; s < NS; s++) F[s] = nmag * (0.5 * (FL[s] + FR[s]) - 0.5 * D[s]);
}

/* src/physics/base/base.cpp:48-62 (+ Roe, new) */
void orc_flux_numerical(int flux, double gamma, const double* UL,
                        const double* UR, const double* N, double* F) {
    switch (flux) {
    case ORC_FLUX_LLF: flux_llf(gamma, UL, UR, N, F); break;
    case ORC_FLUX_HLLC: flux_hllc(gamma, UL, UR, N, F); break;
    case ORC_FLUX_ROE: flux_roe(gamma, UL, UR, N, F); break;
    default: die("no numerical flux function");
    }
}

/* ---- initial conditions / exact solutions ---- */

/* minimum-image coordinate of x relative to a periodic box of length L */
static double wrap(double x, double L) { return x - L * round(x / L); }

static void prim_to_cons(double gam, double rho, const double* u, double p, double* U) {
    U[0] = rho;
    U[1] = rho * u[0]; U[2] = rho * u[1]; U[3] = rho * u[2];
    U[4] = p / (gam - 1.0) + 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / rho;
}

/* Evaluate the IC / exact solution at physical point x and time t.
 *   UNIFORM          src/physics/base/base.cpp:150-158
 *   SPHERE           src/physics/base/base.cpp:177-194
 *   VORTEX3D         src/physics/euler/functions.cpp:147-191 (t ignored)
 *   VORTEX_EXTRUDED  the 2D formula src/physics/euler/functions.cpp:94-145
 *                    applied to (x, y), extruded along z (new); the advected
 *                    core is taken at its minimum periodic image (new);
 *                    optional data[5], data[6]: the core's initial centre
 *                    (default the origin, as the reference's formula).
 *   GAUSSIAN_BUMP    new: rho = 1 + A exp(-(r/W)^2) carried by uniform V, p0.
 *   TAYLOR_GREEN     new: inviscid TGV at Mach M, p0 = 1/(gamma M^2); no
 *                    exact solution, t ignored. */
void orc_ic_eval(const orc_params* prm, const double* x, double t, double* U) {
    const double gam = prm->gamma;
    const double* D = prm->ic_data;
    switch (prm->ic) {
    case ORC_IC_UNIFORM:
        for (int s = 0; s < NS; s++) U[s] = D[s];
        break;
    case ORC_IC_SPHERE: {
        const double v1 = 1., v0 = 0., R = 0.5, w = 0.5;
        double magR = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        for (int s = 0; s < NS; s++) U[s] = 0.5 * (v1 - v0) * (tanh(M_PI * (R - magR) / w) + 1.0) + v0;
        break; }
    case ORC_IC_VORTEX3D: {
        double rho0 = 1., sos0 = 1., L = 1.;
        double Mach = 1.0, radius = L / 10., sigma = 1;
        double beta = Mach * 5. * sqrt(2.) / (4. * M_PI) * exp(0.5);
        double u0 = Mach * sos0;
        double x1 = x[0] - 0.5, y1 = x[2] - 0.5;
        double f = -0.5 / (sigma * sigma) * (x1 * x1 + y1 * y1) / (radius * radius);
        double Om = beta * exp(f);
        double du = -y1 / radius * Om, dv = x1 / radius * Om;
        double dT = -(gam - 1.) / 2. * Om * Om;
        double p = 1. / gam * pow(1 + dT, gam / (gam - 1.));
        U[0] = pow(rho0 + dT, 1. / (gam - 1.));
        U[1] = U[0] * (u0 + du);
        U[2] = 0.;
        U[3] = U[0] * dv;
        double rKE = (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0];
        U[4] = p / (gam - 1.0) + 0.5 * rKE;
        break; }
    case ORC_IC_VORTEX_EXTRUDED: {
        const double rhob = D[0], ub = D[1], vb = D[2], pb = D[3], vs = D[4];
        const double Rg = prm->R;
        if (Rg != 1.0) die("extruded isentropic vortex requires GasConstant = 1");
        double Tb = pb / (rhob * Rg);
        double s = pb / (pow(rhob, gam));
        double xr = wrap((x[0] - D[5]) - ub * t, prm->hi[0] - prm->lo[0]);
        double yr = wrap((x[1] - D[6]) - vb * t, prm->hi[1] - prm->lo[1]);
        double r = sqrt(pow(xr, 2) + pow(yr, 2));
        double dU = vs / (2.0 * M_PI) * exp(0.5 * (1.0 - pow(r, 2)));
        double du = dU * -1.0 * yr, dv = dU * xr;
        double dT = -1.0 * (gam - 1.0) * pow(vs, 2) / (8.0 * gam * pow(M_PI, 2)) * exp(1.0 - pow(r, 2));
        double u = ub + du, v = vb + dv, T = Tb + dT;
        double rho = pow(T / s, 1.0 / (gam - 1.0));
        double rhou = rho * u, rhov = rho * v;
        U[0] = rho; U[1] = rhou; U[2] = rhov; U[3] = 0.;
        U[4] = rho * Rg / (gam - 1.0) * T + 0.5 * (rhou * rhou + rhov * rhov) / rho;
        break; }
    case ORC_IC_GAUSSIAN_BUMP: {
        const double A = D[0], W = D[1];
        const double V[3] = {D[2], D[3], D[4]};
        const double p0 = D[5];
        double r2 = 0.;
        for (int d = 0; d < 3; d++) {
            double xd = wrap(x[d] - D[6 + d] - V[d] * t, prm->hi[d] - prm->lo[d]);
            r2 += xd * xd;
        }
        double rho = 1. + A * exp(-r2 / (W * W));
        prim_to_cons(gam, rho, V, p0, U);
        break; }
    case ORC_IC_TAYLOR_GREEN: {
        const double M = D[0];
        const double p0 = 1. / (gam * M * M);
        double u[3] = {sin(x[0]) * cos(x[1]) * cos(x[2]), -cos(x[0]) * sin(x[1]) * cos(x[2]), 0.};
        double p = p0 + (cos(2. * x[0]) + cos(2. * x[2])) * (cos(2. * x[1]) + 2.) / 16.;
        prim_to_cons(gam, 1., u, p, U);
        break; }
    default: die("unknown initial condition");
    }
}

/* ======================================================================
 * Solver
 * ==================================================================== */

struct orc_solver {
    orc_params prm;
    int p, n, nb;
    long E, F;
    /* mesh (src/mesh/mesh.h:236-302 semantics) */
    long nnodes;
    double* node_coords;     /* [nnodes][3]                */
    long* e2n;               /* [E][8], node n = 4k+2j+i    */
    unsigned* ifaces;        /* [F][8] rankL elemL faceL orientL rankR elemR faceR orientR */
    /* volume helpers at (p+2)^3 */
    int nq;
    double *vq_pts, *vq_wts, *v_bval, *v_bgrad;
    double *djac, *ijac, *vol;
    /* mass matrices from (2p+2)^3 */
    double* imm;             /* [E][nb][nb] */
    /* face helpers */
    int nqf;
    double *fq_wts, *fq_pts;  /* [nqf], [6][nqf][3] */
    double* fq_lp;            /* [nqf][2] face parameters (s,t) of the points */
    double* f_bval;           /* [6*nqf][nb] */
    int *qidxL, *qidxR;       /* [F][nqf] */
    double* normals;          /* [F][nqf][3] */
    /* state */
    double* Uc;
    double time;
    /* residual work arrays (reference shapes) */
    double *Uq, *vgUq, *Uqf, *UqL, *UqR, *Fq, *Fq_elem, *res;
    /* RK arrays */
    double *dU1, *dU2, *dU3, *dU4, *Uhold;
};

static long elem_id(const orc_solver* s, long ix, long iy, long iz) {
    const int* ne = s->prm.nelem;
    ix = (ix % ne[0] + ne[0]) % ne[0];
    iy = (iy % ne[1] + ne[1]) % ne[1];
    iz = (iz % ne[2] + ne[2]) % ne[2];
    return ix + (long)ne[0] * (iy + (long)ne[1] * iz);
}

/* element node coordinates: src/mesh/tools.cpp:6-17 */
static void elem_coords(const orc_solver* s, long e, double* X /* [8][3] */) {
    for (int n = 0; n < 8; ++n)
        for (int d = 0; d < 3; ++d) X[n * 3 + d] = s->node_coords[s->e2n[e * 8 + n] * 3 + d];
}

/* Jacobian at points from gbasis gradients: src/numerics/basis/tools.cpp:42-66
 * (J = X^T dphi via cATxB_to_C, det, closed-form inverse). */
static void element_jacobian(const double* X, int nq, const double* ggrad,
                             double* djac, double* ijac) {
    for (int q = 0; q < nq; ++q) {
        double J[9];
        gemm(3, 3, 8, 1., X, 1, 3, ggrad + (size_t)q * 24, 3, 1, 0., J, 3, 1);
        double d = det3(J);
        if (djac) djac[q] = d;
        if (d <= 0.0) fprintf(stderr, "dg_oracle: negative Jacobian\n");
        if (ijac) inv3(J, ijac + (size_t)q * 9);
    }
}

void orc_box_node(const orc_params* prm, long i, long j, long k, double* x) {
    const int* ne = prm->nelem;
    const long ijk[3] = {i, j, k};
    double h[3], sn[3];
    for (int d = 0; d < 3; ++d) {
        h[d] = (prm->hi[d] - prm->lo[d]) / ne[d];
        sn[d] = sin(2.0 * M_PI * (double)(ijk[d] % ne[d]) / (double)ne[d]);
    }
    for (int d = 0; d < 3; ++d) {
        x[d] = prm->lo[d] + ijk[d] * h[d];
        if (prm->warp != 0.0) x[d] += prm->warp * h[d] * (sn[(d + 1) % 3] * sn[(d + 2) % 3]);
    }
}

/* structured periodic box in the reference mesh conventions
 * (src/mesh/mesh.h:110-156,280-284; faces per shape.h:80-86). */
static void build_mesh(orc_solver* s) {
    const int* ne = s->prm.nelem;
    long nx1 = ne[0] + 1, ny1 = ne[1] + 1, nz1 = ne[2] + 1;
    s->nnodes = nx1 * ny1 * nz1;
    s->node_coords = xcalloc((size_t)s->nnodes * 3, sizeof(double));
    double h[3];
    for (int d = 0; d < 3; ++d) h[d] = (s->prm.hi[d] - s->prm.lo[d]) / ne[d];
    (void)h;
    for (long k = 0; k < nz1; ++k)
        for (long j = 0; j < ny1; ++j)
            for (long i = 0; i < nx1; ++i) {
                long id = i + nx1 * (j + ny1 * k);
                orc_box_node(&s->prm, i, j, k, s->node_coords + id * 3);
            }
    s->e2n = xcalloc((size_t)s->E * 8, sizeof(long));
    for (long iz = 0; iz < ne[2]; ++iz)
        for (long iy = 0; iy < ne[1]; ++iy)
            for (long ix = 0; ix < ne[0]; ++ix) {
                long e = elem_id(s, ix, iy, iz);
                for (int c = 0; c < 2; ++c)
                    for (int b = 0; b < 2; ++b)
                        for (int a = 0; a < 2; ++a)
                            s->e2n[e * 8 + 4 * c + 2 * b + a] =
                                (ix + a) + nx1 * ((iy + b) + ny1 * (iz + c));
            }
}

/* physical location of face points (orientation 0) of element e, face f */
static void face_phys(const orc_solver* s, long e, int f, double* xp /* [nqf][3] */) {
    double X[24];
    elem_coords(s, e, X);
    double* gval = xcalloc((size_t)s->nqf * 8, sizeof(double));
    gbasis_hex(s->nqf, s->fq_pts + (size_t)f * s->nqf * 3, gval, NULL);
    gemm(s->nqf, 3, 8, 1., gval, 8, 1, X, 3, 1, 0., xp, 3, 1);
    free(gval);
}

/* interior face table + orientation lookup (src/solver/helpers.cpp:326-378).
 * orientL = 0; orientR is the unique orientation whose point order makes
 * the paired points coincide physically (modulo the period). */
static void build_faces(orc_solver* s) {
    static const int faceL[3] = {2, 3, 5}, faceR[3] = {4, 1, 0};
    s->F = 3 * s->E;
    s->ifaces = xcalloc((size_t)s->F * 8, sizeof(unsigned));
    s->qidxL = xcalloc((size_t)s->F * s->nqf, sizeof(int));
    s->qidxR = xcalloc((size_t)s->F * s->nqf, sizeof(int));
    const int* ne = s->prm.nelem;
    double per[3];
    for (int d = 0; d < 3; ++d) per[d] = s->prm.hi[d] - s->prm.lo[d];
    /* orientation found once per direction on element 0, checked on all */
    int orientR[3];
    {
        double* xl = xcalloc((size_t)s->nqf * 3, sizeof(double));
        double* xr = xcalloc((size_t)s->nqf * 3, sizeof(double));
        int* po = xcalloc((size_t)s->nqf, sizeof(int));
        for (int d = 0; d < 3; ++d) {
            long e = 0;
            long en = elem_id(s, d == 0, d == 1, d == 2);
            face_phys(s, e, faceL[d], xl);
            face_phys(s, en, faceR[d], xr);
            int found = -1;
            for (int o = 0; o < 8; ++o) {
                orc_face_pts_order(o, s->nqf, po);
                int ok = 1;
                for (int q = 0; q < s->nqf && ok; ++q)
                    for (int c = 0; c < 3; ++c) {
                        double dd = wrap(xl[q * 3 + c] - xr[po[q] * 3 + c], per[c]);
                        if (fabs(dd) > 1e-9 * per[c]) { ok = 0; break; }
                    }
                if (ok) { if (found >= 0) die("ambiguous face orientation"); found = o; }
            }
            if (found < 0) die("no face orientation matches");
            orientR[d] = found;
        }
        free(xl); free(xr); free(po);
    }
    int* po0 = xcalloc((size_t)s->nqf, sizeof(int));
    int* poR[3];
    orc_face_pts_order(0, s->nqf, po0);
    for (int d = 0; d < 3; ++d) {
        poR[d] = xcalloc((size_t)s->nqf, sizeof(int));
        orc_face_pts_order(orientR[d], s->nqf, poR[d]);
    }
    for (long iz = 0; iz < ne[2]; ++iz)
        for (long iy = 0; iy < ne[1]; ++iy)
            for (long ix = 0; ix < ne[0]; ++ix) {
                long e = elem_id(s, ix, iy, iz);
                for (int d = 0; d < 3; ++d) {
                    long f = 3 * e + d;
                    long en = elem_id(s, ix + (d == 0), iy + (d == 1), iz + (d == 2));
                    unsigned* r = s->ifaces + f * 8;
                    r[0] = 0; r[1] = (unsigned)e; r[2] = faceL[d]; r[3] = 0;
                    r[4] = 0; r[5] = (unsigned)en; r[6] = faceR[d]; r[7] = (unsigned)orientR[d];
                    for (int q = 0; q < s->nqf; ++q) {
                        s->qidxL[f * s->nqf + q] = po0[q];
                        s->qidxR[f * s->nqf + q] = poR[d][q];
                    }
                }
            }
    free(po0);
    for (int d = 0; d < 3; ++d) free(poR[d]);
}

/* normals: src/solver/helpers.cpp:380-500 with
 * src/numerics/basis/shape.cpp:482-532; face gbasis gradients at the
 * face-0 reference points (src/numerics/basis/basis.cpp:324-338). */
static void build_normals(orc_solver* s) {
    int nqf = s->nqf;
    s->normals = xcalloc((size_t)s->F * nqf * 3, sizeof(double));
    double xn[2];
    orc_equidistant_nodes(-1., 1., 2, xn);
    #pragma omp parallel for schedule(static)
    for (long f = 0; f < s->F; ++f) {
        /* one rank: the left element is always local, sign = +1 */
        long e = s->ifaces[f * 8 + 1];
        int fid = (int)s->ifaces[f * 8 + 2];
        double sign = 1.0;
        double X[24], fc[12];
        int lf[4];
        elem_coords(s, e, X);
        orc_local_nodes_on_face(fid, 1, lf);
        for (int d = 0; d < 3; ++d)
            for (int n = 0; n < 4; ++n) fc[n * 3 + d] = X[lf[n] * 3 + d];
        for (int q = 0; q < nqf; ++q) {
            /* LagrangeQuad gorder-1 gradients (src/numerics/basis/tools.cpp:
             * 180-213) at the reference's evaluation point, the face-0
             * points' first two coordinates of index q (basis.cpp:324-338);
             * exact_normals: at the left trace's own point qidxL[q] */
            double a = s->fq_pts[q * 3 + 0], bq = s->fq_pts[q * 3 + 1];
            if (s->prm.exact_normals) {
                const int k = s->qidxL[f * nqf + q];
                a = s->fq_lp[k * 2 + 0];
                bq = s->fq_lp[k * 2 + 1];
            }
            double vx[2], vy[2], gx[2], gy[2], fg[8];
            lagrange_val_1d(a, xn, 2, vx); lagrange_val_1d(bq, xn, 2, vy);
            lagrange_grad_1d(a, xn, 2, gx); lagrange_grad_1d(bq, xn, 2, gy);
            for (int j = 0; j < 2; ++j)
                for (int i = 0; i < 2; ++i) {
                    fg[(j * 2 + i) * 2 + 0] = gx[i] * vy[j];
                    fg[(j * 2 + i) * 2 + 1] = vx[i] * gy[j];
                }
            double xg[6];   /* [3][2] = fc^T fg */
            gemm(3, 2, 4, 1., fc, 1, 3, fg, 2, 1, 0., xg, 2, 1);
            double av[3] = {xg[0], xg[2], xg[4]}, b[3] = {xg[1], xg[3], xg[5]};
            double* N = s->normals + ((size_t)f * nqf + q) * 3;
            N[0] = av[1] * b[2] - av[2] * b[1];
            N[1] = av[2] * b[0] - av[0] * b[2];
            N[2] = av[0] * b[1] - av[1] * b[0];
            for (int d = 0; d < 3; ++d) N[d] = sign * N[d];
        }
    }
}

/* inverse mass matrices: src/solver/helpers.cpp:53-100 with
 * src/numerics/basis/tools.cpp:85-110 ((2p+2)^3 points). */
static void build_imm(orc_solver* s) {
    int nb = s->nb;
    int q1 = orc_nq1d_from_qorder(orc_hex_qorder(2 * s->p));
    int nq = q1 * q1 * q1;
    double* pts = xcalloc((size_t)nq * 3, sizeof(double));
    double* wts = xcalloc((size_t)nq, sizeof(double));
    hex_rule(q1, pts, wts);
    double* bval = xcalloc((size_t)nq * nb, sizeof(double));
    double* ggrad = xcalloc((size_t)nq * 24, sizeof(double));
    orc_basis_3d(s->prm.basis, s->p, nq, pts, bval, NULL);
    gbasis_hex(nq, pts, NULL, ggrad);
    s->imm = xcalloc((size_t)s->E * nb * nb, sizeof(double));
    #pragma omp parallel
    {
        double* dj = xcalloc((size_t)nq, sizeof(double));
        double* sb = xcalloc((size_t)nq * nb, sizeof(double));
        double* MM = xcalloc((size_t)nb * nb, sizeof(double));
        #pragma omp for schedule(static)
        for (long e = 0; e < s->E; ++e) {
            double X[24];
            elem_coords(s, e, X);
            element_jacobian(X, nq, ggrad, dj, NULL);
            memcpy(sb, bval, sizeof(double) * nq * nb);
            for (int q = 0; q < nq; ++q) {
                double c = wts[q] * dj[q];
                for (int b = 0; b < nb; ++b) sb[q * nb + b] = c * sb[q * nb + b];
            }
            gemm(nb, nb, nq, 1., bval, 1, nb, sb, nb, 1, 0., MM, nb, 1);
            inv_dense(nb, MM, s->imm + (size_t)e * nb * nb);
        }
        free(dj); free(sb); free(MM);
    }
    free(pts); free(wts); free(bval); free(ggrad);
}

/* volume helpers: src/solver/helpers.cpp:4-51 ((p+2)^3 points) */
static void build_volume_helpers(orc_solver* s) {
    int q1 = orc_nq1d_from_qorder(orc_hex_qorder(s->p));
    s->nq = q1 * q1 * q1;
    int nq = s->nq, nb = s->nb;
    s->vq_pts = xcalloc((size_t)nq * 3, sizeof(double));
    s->vq_wts = xcalloc((size_t)nq, sizeof(double));
    hex_rule(q1, s->vq_pts, s->vq_wts);
    s->v_bval = xcalloc((size_t)nq * nb, sizeof(double));
    s->v_bgrad = xcalloc((size_t)nq * nb * 3, sizeof(double));
    orc_basis_3d(s->prm.basis, s->p, nq, s->vq_pts, s->v_bval, s->v_bgrad);
    double* ggrad = xcalloc((size_t)nq * 24, sizeof(double));
    gbasis_hex(nq, s->vq_pts, NULL, ggrad);
    s->djac = xcalloc((size_t)s->E * nq, sizeof(double));
    s->ijac = xcalloc((size_t)s->E * nq * 9, sizeof(double));
    s->vol = xcalloc((size_t)s->E, sizeof(double));
    #pragma omp parallel for schedule(static)
    for (long e = 0; e < s->E; ++e) {
        double X[24];
        elem_coords(s, e, X);
        element_jacobian(X, nq, ggrad, s->djac + (size_t)e * nq, s->ijac + (size_t)e * nq * 9);
        double v = 0.;   /* src/mesh/tools.cpp:27-36 */
        for (int q = 0; q < nq; ++q) v += s->vq_wts[q] * s->djac[(size_t)e * nq + q];
        s->vol[e] = v;
    }
    free(ggrad);
}

/* face helpers: src/solver/helpers.cpp:205-323 */
static void build_face_helpers(orc_solver* s) {
    int q1 = orc_nq1d_from_qorder(orc_quad_face_qorder(s->p));
    s->nqf = q1 * q1;
    int nqf = s->nqf, nb = s->nb;
    double* lp = xcalloc((size_t)nqf * 2, sizeof(double));
    s->fq_wts = xcalloc((size_t)nqf, sizeof(double));
    quad_rule(q1, lp, s->fq_wts);
    s->fq_pts = xcalloc((size_t)6 * nqf * 3, sizeof(double));
    for (int f = 0; f < 6; ++f) orc_points_on_face(f, 0, nqf, lp, s->fq_pts + (size_t)f * nqf * 3);
    s->f_bval = xcalloc((size_t)6 * nqf * nb, sizeof(double));
    orc_basis_3d(s->prm.basis, s->p, 6 * nqf, s->fq_pts, s->f_bval, NULL);
    s->fq_lp = lp;
}

static void alloc_state(orc_solver* s) {
    size_t E = s->E, nb = s->nb, nq = s->nq, nqf = s->nqf, F = s->F;
    s->Uc = xcalloc(E * nb * NS, sizeof(double));
    s->res = xcalloc(E * nb * NS, sizeof(double));
    s->Uq = xcalloc(E * nq * NS, sizeof(double));
    s->vgUq = xcalloc(E * nq * NS * 3, sizeof(double));
    s->Uqf = xcalloc(E * 6 * nqf * NS, sizeof(double));
    s->UqL = xcalloc(F * nqf * NS, sizeof(double));
    s->UqR = xcalloc(F * nqf * NS, sizeof(double));
    s->Fq = xcalloc(F * nqf * NS, sizeof(double));
    s->Fq_elem = xcalloc(E * 6 * nqf * NS, sizeof(double));
    s->dU1 = xcalloc(E * nb * NS, sizeof(double));
    s->dU2 = xcalloc(E * nb * NS, sizeof(double));
    s->dU3 = xcalloc(E * nb * NS, sizeof(double));
    s->dU4 = xcalloc(E * nb * NS, sizeof(double));
    s->Uhold = xcalloc(E * nb * NS, sizeof(double));
    s->time = 0.0;
}

orc_solver* orc_create(const orc_params* prm) {
    if (prm->order < 1) die("order must be >= 1");
    if (prm->basis == ORC_LEGENDRE_HEX && prm->order > 7) die("LegendreHex supports p <= 7");
    for (int d = 0; d < 3; ++d) if (prm->nelem[d] < 1) die("bad element count");
    orc_solver* s = xcalloc(1, sizeof(orc_solver));
    s->prm = *prm;
    s->p = prm->order; s->n = s->p + 1; s->nb = s->n * s->n * s->n;
    s->E = (long)prm->nelem[0] * prm->nelem[1] * prm->nelem[2];
    build_mesh(s);
    build_imm(s);
    build_volume_helpers(s);
    build_face_helpers(s);
    build_faces(s);
    build_normals(s);
    alloc_state(s);
    return s;
}

/* A mesh given as the reference's mesh datasets (src/mesh/mesh.cpp:79-175):
 * NodeCoords [nnodes][3], Elem2Nodes [E][8] (node 4k + 2j + i), and the
 * IFaceData 6-tuples [F][6] = (elemL, faceL, orientL, elemR, faceR, orientR);
 * one rank.  prm's nelem is not used; lo/hi still give the period of the
 * exact solutions that wrap. */
orc_solver* orc_create_mesh(const orc_params* prm, long nnodes, const double* coords, long E,
                            const long* e2n, long F, const int* iface6) {
    if (prm->order < 1) die("order must be >= 1");
    if (prm->basis == ORC_LEGENDRE_HEX && prm->order > 7) die("LegendreHex supports p <= 7");
    orc_solver* s = xcalloc(1, sizeof(orc_solver));
    s->prm = *prm;
    s->p = prm->order; s->n = s->p + 1; s->nb = s->n * s->n * s->n;
    s->E = E;
    s->nnodes = nnodes;
    s->node_coords = xcalloc((size_t)nnodes * 3, sizeof(double));
    memcpy(s->node_coords, coords, sizeof(double) * nnodes * 3);
    s->e2n = xcalloc((size_t)E * 8, sizeof(long));
    memcpy(s->e2n, e2n, sizeof(long) * E * 8);
    build_imm(s);
    build_volume_helpers(s);
    build_face_helpers(s);
    s->F = F;
    s->ifaces = xcalloc((size_t)F * 8, sizeof(unsigned));
    s->qidxL = xcalloc((size_t)F * s->nqf, sizeof(int));
    s->qidxR = xcalloc((size_t)F * s->nqf, sizeof(int));
    for (long f = 0; f < F; ++f) {
        const int* r = iface6 + f * 6;
        unsigned* t = s->ifaces + f * 8;
        t[0] = 0; t[1] = (unsigned)r[0]; t[2] = (unsigned)r[1]; t[3] = (unsigned)r[2];
        t[4] = 0; t[5] = (unsigned)r[3]; t[6] = (unsigned)r[4]; t[7] = (unsigned)r[5];
        orc_face_pts_order(r[2], s->nqf, s->qidxL + f * s->nqf);
        orc_face_pts_order(r[5], s->nqf, s->qidxR + f * s->nqf);
    }
    build_normals(s);
    alloc_state(s);
    return s;
}

void orc_destroy(orc_solver* s) {
    if (!s) return;
    void* ptrs[] = {s->node_coords, s->e2n, s->ifaces, s->vq_pts, s->vq_wts, s->v_bval,
                    s->v_bgrad, s->djac, s->ijac, s->vol, s->imm, s->fq_wts, s->fq_pts,
                    s->f_bval, s->qidxL, s->qidxR, s->normals, s->fq_lp, s->Uc, s->Uq, s->vgUq,
                    s->Uqf, s->UqL, s->UqR, s->Fq, s->Fq_elem, s->res, s->dU1, s->dU2,
                    s->dU3, s->dU4, s->Uhold};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    free(s);
}

long orc_num_elems(const orc_solver* s) { return s->E; }
int orc_num_basis(const orc_solver* s) { return s->nb; }
long orc_num_ifaces(const orc_solver* s) { return s->F; }
double orc_time(const orc_solver* s) { return s->time; }
void orc_get_face_table(const orc_solver* s, unsigned* faces) {
    memcpy(faces, s->ifaces, sizeof(unsigned) * 8 * s->F);
}
void orc_get_normals(const orc_solver* s, double* normals) {
    memcpy(normals, s->normals, sizeof(double) * 3 * s->nqf * s->F);
}
void orc_get_imm(const orc_solver* s, long e, double* imm) {
    memcpy(imm, s->imm + (size_t)e * s->nb * s->nb, sizeof(double) * s->nb * s->nb);
}
void orc_set_state(orc_solver* s, const double* U) {
    memcpy(s->Uc, U, sizeof(double) * s->E * s->nb * NS);
}
void orc_get_state(const orc_solver* s, double* U) {
    memcpy(U, s->Uc, sizeof(double) * s->E * s->nb * NS);
}
double orc_total_volume(const orc_solver* s) {
    double v = 0.;   /* src/mesh/tools.cpp:38-47 */
    for (long e = 0; e < s->E; ++e) v += s->vol[e];
    return v;
}

/* init_state_from_fcn + L2_projection: src/solver/base.cpp:113-223,
 * src/solver/tools.cpp:66-122 ((2 max(p,1) + 2)^3 points). */
void orc_init_state(orc_solver* s) {
    int nb = s->nb;
    int order_ = s->p > 1 ? s->p : 1;
    int q1 = orc_nq1d_from_qorder(orc_hex_qorder(2 * order_));
    int nq = q1 * q1 * q1;
    double* pts = xcalloc((size_t)nq * 3, sizeof(double));
    double* wts = xcalloc((size_t)nq, sizeof(double));
    hex_rule(q1, pts, wts);
    double* bval = xcalloc((size_t)nq * nb, sizeof(double));
    double* gval = xcalloc((size_t)nq * 8, sizeof(double));
    double* ggrad = xcalloc((size_t)nq * 24, sizeof(double));
    orc_basis_3d(s->prm.basis, s->p, nq, pts, bval, NULL);
    gbasis_hex(nq, pts, gval, ggrad);
    #pragma omp parallel
    {
        double* xp = xcalloc((size_t)nq * 3, sizeof(double));
        double* f = xcalloc((size_t)nq * NS, sizeof(double));
        double* dj = xcalloc((size_t)nq, sizeof(double));
        double* rhs = xcalloc((size_t)nb * NS, sizeof(double));
        #pragma omp for schedule(static)
        for (long e = 0; e < s->E; ++e) {
            double X[24];
            elem_coords(s, e, X);
            gemm(nq, 3, 8, 1., gval, 8, 1, X, 3, 1, 0., xp, 3, 1);
            for (int q = 0; q < nq; ++q) orc_ic_eval(&s->prm, xp + q * 3, s->time, f + q * NS);
            element_jacobian(X, nq, ggrad, dj, NULL);
            for (int q = 0; q < nq; ++q) {
                double c = wts[q] * dj[q];
                for (int v = 0; v < NS; ++v) f[q * NS + v] = c * f[q * NS + v];
            }
            gemm(nb, NS, nq, 1., bval, 1, nb, f, NS, 1, 0., rhs, NS, 1);
            gemm(nb, NS, nb, 1., s->imm + (size_t)e * nb * nb, nb, 1, rhs, NS, 1, 0.,
                 s->Uc + (size_t)e * nb * NS, NS, 1);
        }
        free(xp); free(f); free(dj); free(rhs);
    }
    free(pts); free(wts); free(bval); free(gval); free(ggrad);
}

/* get_residual: src/solver/base.cpp:401-561 */
void orc_residual(orc_solver* s, const double* U, double* R) {
    const int nb = s->nb, nq = s->nq, nqf = s->nqf;
    const long E = s->E, F = s->F;
    const double gam = s->prm.gamma;
    /* fill(res, 0): base.cpp:405-406 */
    memset(R, 0, sizeof(double) * E * nb * NS);
    /* ---- get_element_residuals: base.cpp:415-468 ---- */
    #pragma omp parallel for schedule(static)
    for (long e = 0; e < E; ++e) {
        const double* Ue = U + (size_t)e * nb * NS;
        double* Uq = s->Uq + (size_t)e * nq * NS;
        double* vg = s->vgUq + (size_t)e * nq * NS * 3;
        /* evaluate_state: helpers.cpp:185-198 */
        gemm(nq, NS, nb, 1., s->v_bval, nb, 1, Ue, NS, 1, 0., Uq, NS, 1);
        /* VolumeFluxesFunctor: flux_functors_impl.h:46-115 */
        for (int q = 0; q < nq; ++q) {
            double gF1[15], gF2[15];
            orc_flux_interior(gam, Uq + q * NS, gF1);
            for (int i = 0; i < 15; i++) gF2[i] = 0.;
            const double* ij = s->ijac + ((size_t)e * nq + q) * 9;
            double dj = s->djac[(size_t)e * nq + q], w = s->vq_wts[q];
            for (int id1 = 0; id1 < 3; id1++)
                for (int id2 = 0; id2 < 3; id2++)
                    for (int is = 0; is < NS; is++)
                        gF2[id1 * NS + is] += gF1[id2 * NS + is] * ij[id1 * 3 + id2] * dj * w;
            for (int is = 0; is < NS; is++) Uq[q * NS + is] = 0.;   /* F is unused ({0}) */
            for (int id = 0; id < 3; id++)
                for (int is = 0; is < NS; is++) vg[(q * NS + is) * 3 + id] = gF2[id * NS + is];
        }
        /* calculate_volume_flux_integral: tools.cpp:6-35 */
        double* Re = R + (size_t)e * nb * NS;
        for (int ib = 0; ib < nb; ib++)
            for (int is = 0; is < NS; is++)
                for (int iq = 0; iq < nq; iq++)
                    for (int id = 0; id < 3; id++)
                        Re[ib * NS + is] += s->v_bgrad[((size_t)iq * nb + ib) * 3 + id] *
                                            vg[(iq * NS + is) * 3 + id];
    }
    /* ---- get_interior_face_residuals: base.cpp:470-561 ---- */
    #pragma omp parallel for schedule(static)
    for (long e = 0; e < E; ++e)   /* evaluate_state with face_basis_val */
        gemm(6 * nqf, NS, nb, 1., s->f_bval, nb, 1, U + (size_t)e * nb * NS, NS, 1, 0.,
             s->Uqf + (size_t)e * 6 * nqf * NS, NS, 1);
    #pragma omp parallel for schedule(static)
    for (long f = 0; f < F; ++f) {
        const unsigned* r = s->ifaces + f * 8;
        long eL = r[1], eR = r[5];
        int fL = (int)r[2], fR = (int)r[6];
        for (int q = 0; q < nqf; ++q) {
            /* construct_face_states: base.cpp:287-327 */
            const double* uL = s->Uqf + ((size_t)eL * 6 * nqf + fL * nqf + s->qidxL[f * nqf + q]) * NS;
            const double* uR = s->Uqf + ((size_t)eR * 6 * nqf + fR * nqf + s->qidxR[f * nqf + q]) * NS;
            double* UL = s->UqL + ((size_t)f * nqf + q) * NS;
            double* UR = s->UqR + ((size_t)f * nqf + q) * NS;
            for (int v = 0; v < NS; ++v) { UL[v] = uL[v]; UR[v] = uR[v]; }
            /* InteriorFacesFluxFunctor: flux_functors_impl.h:168-264 */
            double Fh[NS];
            orc_flux_numerical(s->prm.flux, gam, UL, UR, s->normals + ((size_t)f * nqf + q) * 3, Fh);
            for (int v = 0; v < NS; ++v) s->Fq[((size_t)f * nqf + q) * NS + v] = Fh[v] * s->fq_wts[q];
        }
    }
    /* construct_flux_state: base.cpp:329-374 (each slot written once) */
    #pragma omp parallel for schedule(static)
    for (long f = 0; f < F; ++f) {
        const unsigned* r = s->ifaces + f * 8;
        long eL = r[1], eR = r[5];
        int fL = (int)r[2], fR = (int)r[6];
        for (int q = 0; q < nqf; ++q)
            for (int v = 0; v < NS; ++v) {
                double fq = s->Fq[((size_t)f * nqf + q) * NS + v];
                s->Fq_elem[((size_t)eL * 6 * nqf + fL * nqf + s->qidxL[f * nqf + q]) * NS + v] = -1.0 * fq;
                s->Fq_elem[((size_t)eR * 6 * nqf + fR * nqf + s->qidxR[f * nqf + q]) * NS + v] = fq;
            }
    }
    /* calculate_face_flux_integral: tools.cpp:37-63 (res += Bf^T Fq_elem) */
    #pragma omp parallel for schedule(static)
    for (long e = 0; e < E; ++e)
        gemm(nb, NS, 6 * nqf, 1., s->f_bval, 1, nb, s->Fq_elem + (size_t)e * 6 * nqf * NS, NS, 1,
             1., R + (size_t)e * nb * NS, NS, 1);
}

/* mult_inv_mass_matrix: src/solver/tools.cpp:124-139 (dU = dt iMM res) */
void orc_inv_mass(const orc_solver* s, double dt, const double* R, double* dU) {
    const int nb = s->nb;
    #pragma omp parallel for schedule(static)
    for (long e = 0; e < s->E; ++e)
        gemm(nb, NS, nb, dt, s->imm + (size_t)e * nb * nb, nb, 1, R + (size_t)e * nb * NS, NS, 1,
             0., dU + (size_t)e * nb * NS, NS, 1);
}

/* KokkosBlas::update(c, A, 1, B, 0, C): C = c A + B  (linear_algebra.cpp:83-90) */
static void capb_to_c(size_t n, double c, const double* A, const double* B, double* C) {
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) C[i] = c * A[i] + B[i];
}
/* KokkosBlas::axpy(c, A, B): B = B + c A  (linear_algebra.cpp:75-81) */
static void capb_to_b(size_t n, double c, const double* A, double* B) {
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) B[i] = B[i] + c * A[i];
}

/* FE: src/numerics/timestepping/stepper.cpp:23-55;
 * RK4: src/numerics/timestepping/stepper.cpp:71-142 */
void orc_step(orc_solver* s, int stepper, double dt, int nsteps) {
    size_t n = (size_t)s->E * s->nb * NS;
    for (int it = 0; it < nsteps; ++it) {
        if (stepper == ORC_FE) {
            orc_residual(s, s->Uc, s->res);
            orc_inv_mass(s, dt, s->res, s->dU1);
            capb_to_b(n, 1., s->dU1, s->Uc);
            s->time += dt;
        } else {
            memcpy(s->Uhold, s->Uc, sizeof(double) * n);
            orc_residual(s, s->Uc, s->res);
            orc_inv_mass(s, dt, s->res, s->dU1);
            capb_to_c(n, 0.5, s->dU1, s->Uhold, s->Uc);
            s->time += dt / 2.0;
            orc_residual(s, s->Uc, s->res);
            orc_inv_mass(s, dt, s->res, s->dU2);
            capb_to_c(n, 0.5, s->dU2, s->Uhold, s->Uc);
            orc_residual(s, s->Uc, s->res);
            orc_inv_mass(s, dt, s->res, s->dU3);
            capb_to_c(n, 1.0, s->dU3, s->Uhold, s->Uc);
            s->time += dt / 2.0;
            orc_residual(s, s->Uc, s->res);
            orc_inv_mass(s, dt, s->res, s->dU4);
            capb_to_b(n, 2.0, s->dU3, s->dU4);
            capb_to_b(n, 2.0, s->dU2, s->dU4);
            capb_to_b(n, 1.0, s->dU1, s->dU4);
            capb_to_b(n, 1.0 / 6.0, s->dU4, s->Uhold);
            memcpy(s->Uc, s->Uhold, sizeof(double) * n);
        }
    }
}

void orc_flux_interior_n(double gamma, int n, const double* U, double* Fdir) {
    for (int i = 0; i < n; ++i) orc_flux_interior(gamma, U + (size_t)i * NS, Fdir + (size_t)i * 3 * NS);
}

void orc_flux_numerical_n(int flux, double gamma, int n, const double* UL, const double* UR,
                          const double* N, double* F) {
    for (int i = 0; i < n; ++i)
        orc_flux_numerical(flux, gamma, UL + (size_t)i * NS, UR + (size_t)i * NS, N + (size_t)i * 3,
                           F + (size_t)i * NS);
}

/* get_error: src/post/main.cpp:30-201 (ord = 2, normalised by volume) */
void orc_l2_error(const orc_solver* s, double t, double* err) {
    int nb = s->nb;
    int order_ = s->p > 1 ? s->p : 1;
    int q1 = orc_nq1d_from_qorder(orc_hex_qorder(2 * order_));
    int nq = q1 * q1 * q1;
    double* pts = xcalloc((size_t)nq * 3, sizeof(double));
    double* wts = xcalloc((size_t)nq, sizeof(double));
    hex_rule(q1, pts, wts);
    double* bval = xcalloc((size_t)nq * nb, sizeof(double));
    double* gval = xcalloc((size_t)nq * 8, sizeof(double));
    double* ggrad = xcalloc((size_t)nq * 24, sizeof(double));
    orc_basis_3d(s->prm.basis, s->p, nq, pts, bval, NULL);
    gbasis_hex(nq, pts, gval, ggrad);
    double* eel = xcalloc((size_t)s->E * NS, sizeof(double));
    #pragma omp parallel
    {
        double* xp = xcalloc((size_t)nq * 3, sizeof(double));
        double* ue = xcalloc((size_t)nq * NS, sizeof(double));
        double* uq = xcalloc((size_t)nq * NS, sizeof(double));
        double* dj = xcalloc((size_t)nq, sizeof(double));
        #pragma omp for schedule(static)
        for (long e = 0; e < s->E; ++e) {
            double X[24];
            elem_coords(s, e, X);
            gemm(nq, NS, nb, 1., bval, nb, 1, s->Uc + (size_t)e * nb * NS, NS, 1, 0., uq, NS, 1);
            gemm(nq, 3, 8, 1., gval, 8, 1, X, 3, 1, 0., xp, 3, 1);
            for (int q = 0; q < nq; ++q) orc_ic_eval(&s->prm, xp + q * 3, t, ue + q * NS);
            element_jacobian(X, nq, ggrad, dj, NULL);
            for (int v = 0; v < NS; ++v) {
                double acc = 0.;
                for (int q = 0; q < nq; ++q)
                    acc += pow(uq[q * NS + v] - ue[q * NS + v], 2) * wts[q] * dj[q];
                eel[e * NS + v] = acc;
            }
        }
        free(xp); free(ue); free(uq); free(dj);
    }
    double tv = orc_total_volume(s);
    for (int v = 0; v < NS; ++v) {
        double tot = 0.;
        for (long e = 0; e < s->E; ++e) tot += eel[e * NS + v];
        err[v] = pow(tot / tv, 1. / 2);
    }
    free(eel); free(pts); free(wts); free(bval); free(gval); free(ggrad);
}

/* max |u|+c over the volume quadrature points of the current state
 * (the wave speed of src/physics/euler/functions.cpp:25-40; used by the
 * CFL time-step rule, which the reference does not have). */
double orc_max_wavespeed_state(const orc_solver* s) {
    int nb = s->nb, nq = s->nq;
    double amax = 0.;
    #pragma omp parallel
    {
        double* uq = xcalloc((size_t)nq * NS, sizeof(double));
        double loc = 0.;
        #pragma omp for schedule(static)
        for (long e = 0; e < s->E; ++e) {
            gemm(nq, NS, nb, 1., s->v_bval, nb, 1, s->Uc + (size_t)e * nb * NS, NS, 1, 0., uq, NS, 1);
            for (int q = 0; q < nq; ++q) {
                double a = orc_maxwavespeed(s->prm.gamma, uq + q * NS);
                if (a > loc) loc = a;
            }
        }
        #pragma omp critical
        if (loc > amax) amax = loc;
        free(uq);
    }
    return amax;
}

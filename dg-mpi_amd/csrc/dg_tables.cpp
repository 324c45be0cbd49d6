// Reference-element tables of the tensor-product hex bases.
#include "dg_tables.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <stdexcept>

namespace dg {

namespace {
using q128 = __float128;

q128 qabs(q128 x) { return x < 0 ? -x : x; }

// P_n(z) and P_n'(z) by the three-term recurrence, binary128.
void legendre_q(int n, q128 z, q128& pn, q128& dpn) {
    q128 p0 = 1, p1 = z;
    for (int k = 2; k <= n; ++k) {
        q128 p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    pn = n == 0 ? q128(1) : p1;
    dpn = n * (z * p1 - p0) / (z * z - 1);
}
}  // namespace

void gl_rule(int n, double* x, double* w) {
    if (n < 1) throw std::invalid_argument("gl_rule: n < 1");
    if (n == 1) {
        x[0] = 0.0;
        w[0] = 2.0;
        return;
    }
    for (int i = 0; i < n / 2; ++i) {
        // Newton from the Tricomi-style initial guess; converged roots are
        // refined in binary128 then rounded once to double
        q128 z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        q128 pn, dp;
        for (int it = 0; it < 100; ++it) {
            legendre_q(n, z, pn, dp);
            q128 dz = pn / dp;
            z -= dz;
            if (qabs(dz) < q128(1e-33)) break;
        }
        legendre_q(n, z, pn, dp);
        q128 wq = 2 / ((1 - z * z) * dp * dp);
        x[n - 1 - i] = (double)z;
        x[i] = -(double)z;
        w[n - 1 - i] = (double)wq;
        w[i] = (double)wq;
    }
    if (n % 2 == 1) {
        q128 pn, dp;
        legendre_q(n, 0, pn, dp);
        x[n / 2] = 0.0;
        w[n / 2] = (double)(2 / (dp * dp));
    }
}

void basis_1d(Basis b, int p, double x, double* val, double* grad) {
    if (b == Basis::LagrangeHex) {
        double xn[16];
        const int nn = p + 1;
        const double dx = (1. - -1.) / ((double)nn - 1.);
        for (int i = 0; i < nn; i++) xn[i] = -1. + (double)i * dx;
        for (int i = 0; i < nn; i++) {
            double v = 1.;
            for (int j = 0; j < nn; j++)
                if (j != i) v *= (x - xn[j]) / (xn[i] - xn[j]);
            val[i] = v;
            double g = 0.;
            for (int j = 0; j < nn; j++) {
                if (j == i) continue;
                double prod = 1.;
                for (int k = 0; k < nn; k++) {
                    if (k == i) continue;
                    prod *= (k == j) ? 1. / (xn[i] - xn[k]) : (x - xn[k]) / (xn[i] - xn[k]);
                }
                g += prod;
            }
            grad[i] = g;
        }
        return;
    }
    // unnormalised Legendre polynomials and derivatives, p <= 7
    const double x2 = x * x, x3 = x2 * x, x4 = x3 * x, x5 = x4 * x, x6 = x5 * x, x7 = x6 * x;
    const double v[8] = {1., x, 0.5 * (3. * x2 - 1.), 0.5 * (5. * x3 - 3. * x),
                         0.125 * (35. * x4 - 30. * x2 + 3.), 0.125 * (63. * x5 - 70. * x3 + 15. * x),
                         0.0625 * (231. * x6 - 315. * x4 + 105. * x2 - 5.),
                         0.0625 * (429. * x7 - 693. * x5 + 315. * x3 - 35. * x)};
    const double g[8] = {0., 1., 3. * x, 0.5 * (15. * x2 - 3.), 0.125 * (140. * x3 - 60. * x),
                         0.125 * (315. * x4 - 210. * x2 + 15.),
                         0.0625 * (1386. * x5 - 1260. * x3 + 210. * x),
                         0.0625 * (3003. * x6 - 3465. * x4 + 945. * x2 - 35.)};
    for (int i = 0; i <= p; ++i) {
        val[i] = v[i];
        grad[i] = g[i];
    }
}

// bf16 of a float, round to nearest even, and back
static uint16_t bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static float bf16_value(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

KTables make_tables(const CaseConfig& c) {
    const int p = c.order, n = p + 1, m = p + 2, q = 2 * p + 2;
    if (p < 1 || p > PMAX) throw UnsupportedError("order p = " + std::to_string(p) + " has no GPU kernel (1.." +
                                                  std::to_string(PMAX) + ")");
    KTables t;
    std::memset(&t, 0, sizeof(t));
    double x[32], w[32], v[16], g[16];
    // (p+2)-point rule: volume and face quadrature (quadrature/tools.cpp:7-24,
    // quadrilateral.cpp:100-105)
    gl_rule(m, x, w);
    for (int a = 0; a < m; ++a) t.xv[a] = x[a];
    for (int a = 0; a < m; ++a) {
        basis_1d(c.basis, p, x[a], v, g);
        for (int i = 0; i < n; ++i) {
            t.B[a][i] = v[i];
            t.D[a][i] = g[i];
        }
        t.w[a] = w[a];
    }
    // (p+1)-point segment rule: the face quadrature of quadrilaterals
    // (solver/helpers.cpp:225, quadrature/segment.cpp:424-426: order 2p+1)
    gl_rule(n, x, w);
    for (int a = 0; a < n; ++a) {
        basis_1d(c.basis, p, x[a], v, g);
        for (int i = 0; i < n; ++i) t.BF[a][i] = v[i];
        t.wf[a] = w[a];
    }
    basis_1d(c.basis, p, -1.0, v, g);
    for (int i = 0; i < n; ++i) t.phim[i] = v[i];
    basis_1d(c.basis, p, 1.0, v, g);
    for (int i = 0; i < n; ++i) t.phip[i] = v[i];
    // (2p+2)-point rule: mass matrix, IC projection, L2 error
    gl_rule(q, x, w);
    for (int a = 0; a < q; ++a) {
        basis_1d(c.basis, p, x[a], v, g);
        for (int i = 0; i < n; ++i) t.BQ[a][i] = v[i];
        t.xq[a] = x[a];
        t.wq[a] = w[a];
    }
    // 1D mass matrix on the (2p+2)-point rule (exact for degree 2p) and its
    // inverse by Gauss-Jordan elimination in binary128
    q128 Mq[8][16];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < 2 * n; ++j) Mq[i][j] = 0;
    for (int a = 0; a < q; ++a)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) Mq[i][j] += (q128)t.BQ[a][i] * (q128)t.BQ[a][j] * (q128)w[a];
    for (int i = 0; i < n; ++i) Mq[i][n + i] = 1;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (qabs(Mq[r][col]) > qabs(Mq[piv][col])) piv = r;
        if (piv != col)
            for (int j = 0; j < 2 * n; ++j) { q128 tt = Mq[col][j]; Mq[col][j] = Mq[piv][j]; Mq[piv][j] = tt; }
        q128 d = Mq[col][col];
        for (int j = 0; j < 2 * n; ++j) Mq[col][j] /= d;
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            q128 f = Mq[r][col];
            for (int j = 0; j < 2 * n; ++j) Mq[r][j] -= f * Mq[col][j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) t.minv[i][j] = (double)Mq[i][n + j];
    // coefficients of the constant function 1: M^-1 (integral of phi_j),
    // in binary128 (1, 0, ... for Legendre; all 1 for Lagrange)
    for (int i = 0; i < n; ++i) {
        q128 acc = 0;
        for (int j = 0; j < n; ++j) {
            q128 ij = 0;
            for (int a = 0; a < q; ++a) ij += (q128)t.BQ[a][j] * (q128)w[a];
            acc += Mq[i][n + j] * ij;
        }
        t.one[i] = (double)acc;
    }
    {
        bool delta = std::fabs(t.one[0] - 1.0) < 1e-14, ones = true;
        for (int i = 0; i < n; ++i) {
            if (i > 0 && std::fabs(t.one[i]) >= 1e-14) delta = false;
            if (std::fabs(t.one[i] - 1.0) >= 1e-14) ones = false;
        }
        t.one_kind = delta ? 0 : (ones ? 1 : 2);
    }
    t.minv_diag = c.basis == Basis::LegendreHex ? 1 : 0;
    if (t.minv_diag)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                if (i != j) t.minv[i][j] = 0.0;
    for (int a = 0; a < PMAX + 2; ++a) {
        for (int i = 0; i < PMAX + 1; ++i) {
            t.B32[a][i] = (float)t.B[a][i];
            t.D32[a][i] = (float)t.D[a][i];
            t.Bw[a][i] = t.w[a] * t.B[a][i];
            t.Dw[a][i] = t.w[a] * t.D[a][i];
            t.Bw32[a][i] = (float)t.Bw[a][i];
            t.Dw32[a][i] = (float)t.Dw[a][i];
        }
        t.w32[a] = (float)t.w[a];
    }
    for (int i = 0; i < PMAX + 1; ++i) {
        t.phim32[i] = (float)t.phim[i];
        t.phip32[i] = (float)t.phip[i];
    }
    // x-y projection matrix of the bf16 form (zero padding from the memset)
    const int mm = m * m;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
            for (int dir = 0; dir < 3; ++dir)
                for (int b = 0; b < m; ++b)
                    for (int a = 0; a < m; ++a) {
                        const double v = dir == 0 ? t.Dw[a][i] * t.Bw[b][j]
                                                  : (dir == 1 ? t.Bw[a][i] * t.Dw[b][j] : t.Bw[a][i] * t.Bw[b][j]);
                        const float f = (float)v;
                        const uint16_t hi = bf16_rne(f);
                        t.Kproj[0][j * n + i][dir * mm + b * m + a] = hi;
                        t.Kproj[1][j * n + i][dir * mm + b * m + a] = bf16_rne(f - bf16_value(hi));
                    }
    return t;
}

KGeom make_geometry(const CaseConfig& c) {
    KGeom g;
    std::memset(&g, 0, sizeof(g));
    long E = 1;
    for (int d = 0; d < 3; ++d) {
        g.ne[d] = c.nelem[d];
        g.lo[d] = c.lower[d];
        g.period[d] = c.upper[d] - c.lower[d];
        g.h[d] = g.period[d] / c.nelem[d];
        E *= c.nelem[d];
    }
    g.E = (unsigned)E;
    g.dim = c.dim;
    if (c.dim == 2) {
        // quadrilaterals: |J| = hx hy / 4, |N| = half the edge length
        g.jdet = (0.5 * g.h[0]) * (0.5 * g.h[1]);
        g.jinv = 1.0 / g.jdet;
        g.gvol[0] = (2.0 / g.h[0]) * g.jdet;
        g.gvol[1] = (2.0 / g.h[1]) * g.jdet;
        g.farea[0] = 0.5 * g.h[1];
        g.farea[1] = 0.5 * g.h[0];
    } else {
        g.jdet = (0.5 * g.h[0]) * (0.5 * g.h[1]) * (0.5 * g.h[2]);
        g.jinv = 1.0 / g.jdet;
        for (int d = 0; d < 3; ++d) {
            g.gvol[d] = (2.0 / g.h[d]) * g.jdet;
            const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
            g.farea[d] = (0.5 * g.h[t1]) * (0.5 * g.h[t2]);
        }
    }
    g.gamma = c.gamma;
    g.R = c.gas_constant;
    g.mapped = (c.warp > 0.0 || c.mesh_input) ? 1 : 0;
    g.fp32 = c.fp32 ? 1 : 0;
    g.proj = c.proj;
    g.exact_normals = c.exact_normals ? 1 : 0;
    g.xcd_remap = (E % 8 == 0 && E >= 8 * 64) ? 1 : 0;
    // Stage-kernel traversal in bricks of up to 8^3 elements (edge: the
    // largest power of two <= 8 dividing the element count), so that the
    // neighbours of the elements in flight on an XCD are still in its L2.
    // DG_ELEMENT_TILE="lx,ly,lz" (log2 edges) overrides, for sweeps.
    for (int d = 0; d < 3; ++d) {
        int l = 0;
        while (l < 3 && c.nelem[d] % (2 << l) == 0) ++l;
        g.tile[d] = l;
    }
    if (const char* env = std::getenv("DG_ELEMENT_TILE")) {
        int t[3];
        if (std::sscanf(env, "%d,%d,%d", &t[0], &t[1], &t[2]) == 3) {
            bool ok = true;
            for (int d = 0; d < 3; ++d) ok = ok && t[d] >= 0 && t[d] < 16 && c.nelem[d] % (1 << t[d]) == 0;
            if (ok)
                for (int d = 0; d < 3; ++d) g.tile[d] = t[d];
        }
    }
    // a given mesh is traversed in its own element order
    if (c.mesh_input) g.tile[0] = g.tile[1] = g.tile[2] = 0;
    if (c.dim == 2) {
        g.tile[0] = g.tile[1] = g.tile[2] = 0;
        g.xcd_remap = 0;
    }
    return g;
}

void box_node(const CaseConfig& c, long i, long j, long k, double* x) {
    const long ijk[3] = {i, j, k};
    double h[3], sn[3];
    for (int d = 0; d < 3; ++d) {
        h[d] = (c.upper[d] - c.lower[d]) / c.nelem[d];
        sn[d] = std::sin(2.0 * M_PI * (double)(ijk[d] % c.nelem[d]) / (double)c.nelem[d]);
    }
    for (int d = 0; d < 3; ++d) {
        x[d] = c.lower[d] + ijk[d] * h[d];
        if (c.warp != 0.0) x[d] += c.warp * h[d] * (sn[(d + 1) % 3] * sn[(d + 2) % 3]);
    }
}

std::vector<double> element_nodes(const CaseConfig& c) {
    const long nx = c.nelem[0], ny = c.nelem[1], nz = c.nelem[2];
    std::vector<double> X((size_t)nx * ny * nz * 24);
    for (long iz = 0; iz < nz; ++iz)
        for (long iy = 0; iy < ny; ++iy)
            for (long ix = 0; ix < nx; ++ix) {
                double* xe = X.data() + (size_t)(ix + nx * (iy + ny * iz)) * 24;
                for (int n = 0; n < 8; ++n) box_node(c, ix + (n & 1), iy + ((n >> 1) & 1), iz + (n >> 2), xe + 3 * n);
            }
    return X;
}

// ---------------------------------------------------------------------------
// Face conventions of the reference hexahedron (src/numerics/basis/
// shape.cpp:236-428), restated for the face records of the general path
// ---------------------------------------------------------------------------
namespace {

const int kFaceDir[6] = {2, 1, 0, 1, 0, 2};    // faces 0 z-, 1 y-, 2 x+, 3 y+, 4 x-, 5 z+
const int kFaceSide[6] = {0, 0, 1, 1, 0, 1};
// corner nodes of each face in the reference's cyclic order
const int kFaceNodes[6][4] = {{0, 2, 3, 1}, {0, 1, 5, 4}, {1, 3, 7, 5}, {3, 2, 6, 7}, {2, 0, 4, 6}, {4, 5, 7, 6}};

int face_id(int d, int side) {
    return d == 0 ? (side ? 2 : 4) : (d == 1 ? (side ? 3 : 1) : (side ? 5 : 0));
}

void tangential(int d, int& t1, int& t2) {
    t1 = d == 0 ? 1 : 0;
    t2 = d == 2 ? 1 : 2;
}

// element coordinates of face parameter (s, t), orientation 0: bilinear
// over the face's corners in cyclic order (shape.cpp:296-323)
void face_map(int f, double s, double t, double* xi) {
    const double xb = 0.5 * (s + 1.0), yb = 0.5 * (t + 1.0);
    const double b[4] = {(1. - xb) * (1. - yb), xb * (1. - yb), xb * yb, (1. - xb) * yb};
    for (int r = 0; r < 3; ++r) {
        double v = 0.;
        for (int j = 0; j < 4; ++j) {
            const int n = kFaceNodes[f][j];
            const double c = r == 0 ? 2. * (n & 1) - 1. : (r == 1 ? 2. * ((n >> 1) & 1) - 1. : 2. * (n >> 2) - 1.);
            v += b[j] * c;
        }
        xi[r] = v;
    }
}

// the eight face point orderings (shape.cpp:326-428)
void face_pts_order(int orient, int npts, int* po) {
    const int n = (int)std::lround(std::sqrt((double)npts)), np1 = npts + 1;
    switch (orient) {
    case 0:
        for (int i = 0; i < npts; i++) po[i] = i;
        break;
    case 1: {
        int c = n - 1;
        for (int i = 0; i < npts; i++) {
            po[i] = c;
            c = (c + n) % np1;
        }
        break;
    }
    case 2:
        for (int i = 0; i < npts; i++) po[i] = npts - 1 - i;
        break;
    case 3: {
        int c = npts - n;
        for (int i = 0; i < npts; i++) {
            po[i] = c;
            c -= n;
            if (c < 0) c += np1;
        }
        break;
    }
    case 4: {
        int c = 0;
        for (int i = 0; i < npts; i++) {
            po[i] = c;
            c += n;
            if (c >= npts) c -= npts - 1;
        }
        break;
    }
    case 5: {
        int c = n - 1, k = 0;
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < n; j++) po[k++] = c--;
            c += 2 * n;
        }
        break;
    }
    case 6: {
        int c = npts - 1;
        for (int i = 0; i < npts; i++) {
            po[i] = c;
            c -= n;
            if (c < 0) c += npts - 1;
        }
        break;
    }
    case 7: {
        int c = npts - n, k = 0;
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < n; j++) po[k++] = c++;
            c -= 2 * n;
        }
        break;
    }
    default:
        throw InputError("face orientation " + std::to_string(orient) + " is not 0..7");
    }
}

// face f's grid point (a1, a2) (xi_t1 = x[a1], xi_t2 = x[a2]) -> canonical
// point index k = j m + i of parameters (x[i], x[j]); and the inverse
struct FaceGrid {
    std::vector<int> g2c, c2g;   // [m*m]
};

FaceGrid face_grid(int f, int m, const double* x) {
    FaceGrid g;
    g.g2c.assign(m * m, -1);
    g.c2g.assign(m * m, -1);
    int t1, t2;
    tangential(kFaceDir[f], t1, t2);
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
            double xi[3];
            face_map(f, x[i], x[j], xi);
            int a1 = -1, a2 = -1;
            for (int a = 0; a < m; ++a) {
                if (std::fabs(xi[t1] - x[a]) < 1e-12) a1 = a;
                if (std::fabs(xi[t2] - x[a]) < 1e-12) a2 = a;
            }
            if (a1 < 0 || a2 < 0) throw std::logic_error("face grid: point not on the tensor grid");
            g.g2c[a2 * m + a1] = j * m + i;
            g.c2g[j * m + i] = a2 * m + a1;
        }
    return g;
}

int map_code(const std::vector<int>& b_of_a, int m) {
    for (int code = 0; code < 8; ++code) {
        bool ok = true;
        for (int a2 = 0; a2 < m && ok; ++a2)
            for (int a1 = 0; a1 < m && ok; ++a1) {
                const int c1 = (code & 1) ? a2 : a1, c2 = (code & 1) ? a1 : a2;
                const int b1 = (code & 2) ? m - 1 - c1 : c1, b2 = (code & 4) ? m - 1 - c2 : c2;
                ok = b_of_a[a2 * m + a1] == b2 * m + b1;
            }
        if (ok) return code;
    }
    throw InputError("IFaceData: a face pairing is not a symmetry of the face point grid");
}

std::vector<double> normal_points(const KTables& t, int m, bool exact) {
    std::vector<double> nev((size_t)48 * m * m * 2);
    std::vector<int> po(m * m), inv(m * m);
    for (int fL = 0; fL < 6; ++fL) {
        const FaceGrid g = face_grid(fL, m, t.xv);
        int t1, t2;
        tangential(kFaceDir[fL], t1, t2);
        for (int oL = 0; oL < 8; ++oL) {
            face_pts_order(oL, m * m, po.data());
            for (int q = 0; q < m * m; ++q) inv[po[q]] = q;
            for (int c2 = 0; c2 < m; ++c2)
                for (int c1 = 0; c1 < m; ++c1) {
                    double* o = nev.data() + (((size_t)(fL * 8 + oL) * m * m) + c2 * m + c1) * 2;
                    if (exact) {
                        o[0] = t.xv[c1];
                        o[1] = t.xv[c2];
                        continue;
                    }
                    // the flux point's index q on the face; the reference
                    // evaluates the normal at the face-0 points' first two
                    // coordinates of index q, (t_q, s_q), as (s, t) of face fL
                    const int q = inv[g.g2c[c2 * m + c1]];
                    const double sq = t.xv[q % m], tq = t.xv[q / m];
                    double xi[3];
                    face_map(fL, tq, sq, xi);
                    o[0] = xi[t1];
                    o[1] = xi[t2];
                }
        }
    }
    return nev;
}

}  // namespace

FaceTables box_face_tables(const CaseConfig& c) {
    FaceTables ft;
    const KTables t = make_tables(c);
    const int m = c.order + 2;
    const long nx = c.nelem[0], ny = c.nelem[1], nz = c.nelem[2], E = nx * ny * nz;
    ft.frec.resize((size_t)E * 12);
    for (long iz = 0; iz < nz; ++iz)
        for (long iy = 0; iy < ny; ++iy)
            for (long ix = 0; ix < nx; ++ix) {
                const long e = ix + nx * (iy + ny * iz);
                for (int d = 0; d < 3; ++d)
                    for (int side = 0; side < 2; ++side) {
                        long cc[3] = {ix, iy, iz};
                        cc[d] = (cc[d] + (side ? 1 : -1) + c.nelem[d]) % c.nelem[d];
                        const long en = cc[0] + nx * (cc[1] + ny * cc[2]);
                        const unsigned meta = (unsigned)face_id(d, 1 - side) | (0u << 3) | ((unsigned)side << 6) |
                                              ((unsigned)(face_id(d, 1) * 8) << 8);
                        ft.frec[((size_t)e * 6 + d * 2 + side) * 2] = (unsigned)en;
                        ft.frec[((size_t)e * 6 + d * 2 + side) * 2 + 1] = meta;
                    }
            }
    ft.nev = normal_points(t, m, c.exact_normals);
    return ft;
}

FaceTables mesh_face_tables(const CaseConfig& c, long E, long F, const int* iface6) {
    FaceTables ft;
    const KTables t = make_tables(c);
    const int m = c.order + 2;
    std::vector<FaceGrid> grid;
    for (int f = 0; f < 6; ++f) grid.push_back(face_grid(f, m, t.xv));
    ft.frec.assign((size_t)E * 12, 0);
    std::vector<char> seen((size_t)E * 6, 0);
    std::vector<int> poL(m * m), poR(m * m), invL(m * m), invR(m * m), b_of_a(m * m), a_of_b(m * m);
    for (long f = 0; f < F; ++f) {
        const int* r = iface6 + f * 6;
        const long eL = r[0], eR = r[3];
        const int fL = r[1], oL = r[2], fR = r[4], oR = r[5];
        if (eL < 0 || eL >= E || eR < 0 || eR >= E || fL < 0 || fL > 5 || fR < 0 || fR > 5)
            throw InputError("IFaceData row " + std::to_string(f) + " is out of range");
        face_pts_order(oL, m * m, poL.data());
        face_pts_order(oR, m * m, poR.data());
        for (int q = 0; q < m * m; ++q) {
            invL[poL[q]] = q;
            invR[poR[q]] = q;
        }
        // L grid point -> canonical L point -> face point q -> canonical R
        // point -> R grid point (the pairing of construct_face_states,
        // src/solver/base.cpp:287-327)
        for (int g = 0; g < m * m; ++g) {
            const int q = invL[grid[fL].g2c[g]];
            b_of_a[g] = grid[fR].c2g[poR[q]];
            const int qq = invR[grid[fR].g2c[g]];
            a_of_b[g] = grid[fL].c2g[poL[qq]];
        }
        const int sL = kFaceDir[fL] * 2 + kFaceSide[fL], sR = kFaceDir[fR] * 2 + kFaceSide[fR];
        if (seen[(size_t)eL * 6 + sL]++ || seen[(size_t)eR * 6 + sR]++)
            throw InputError("IFaceData: an element face appears twice");
        const unsigned lfo = (unsigned)(fL * 8 + oL) << 8;
        ft.frec[((size_t)eL * 6 + sL) * 2] = (unsigned)eR;
        ft.frec[((size_t)eL * 6 + sL) * 2 + 1] = (unsigned)fR | ((unsigned)map_code(b_of_a, m) << 3) | (1u << 6) | lfo;
        ft.frec[((size_t)eR * 6 + sR) * 2] = (unsigned)eL;
        ft.frec[((size_t)eR * 6 + sR) * 2 + 1] = (unsigned)fL | ((unsigned)map_code(a_of_b, m) << 3) | lfo;
    }
    for (size_t i = 0; i < seen.size(); ++i)
        if (!seen[i])
            throw UnsupportedError("element " + std::to_string(i / 6) +
                                   " has a face without a neighbour: boundary faces are not on this path "
                                   "(periodic meshes only)");
    ft.nev = normal_points(t, m, c.exact_normals);
    return ft;
}

FaceTables mesh_face_tables_2d(const CaseConfig& c, long E, long F, const int* iface6) {
    (void)c;
    // quadrilateral faces (shape.cpp:69-122): 0 (n0 n1, y-), 1 (n1 n3, x+),
    // 2 (n3 n2, y+), 3 (n2 n0, x-); slot d*2 + side of the face's normal
    // direction; sigma: +1 when the face parameter runs along the element's
    // tangential coordinate, -1 against it; orientation 1 reverses it
    auto dir = [](int f) { return (f == 1 || f == 3) ? 0 : 1; };
    auto side = [](int f) { return (f == 1 || f == 2) ? 1 : 0; };
    auto sigma = [](int f, int o) { return ((f <= 1) ? 1 : -1) * (o ? -1 : 1); };
    FaceTables ft;
    ft.frec.assign((size_t)E * 8, 0);
    std::vector<char> seen((size_t)E * 4, 0);
    for (long f = 0; f < F; ++f) {
        const int* r = iface6 + f * 6;
        const long eL = r[0], eR = r[3];
        const int fL = r[1], oL = r[2], fR = r[4], oR = r[5];
        if (eL < 0 || eL >= E || eR < 0 || eR >= E || fL < 0 || fL > 3 || fR < 0 || fR > 3)
            throw InputError("IFaceData row " + std::to_string(f) + " is out of range");
        if (oL < 0 || oL > 1 || oR < 0 || oR > 1)
            throw InputError("IFaceData row " + std::to_string(f) + ": face orientation is not 0..1");
        const int sL = dir(fL) * 2 + side(fL), sR = dir(fR) * 2 + side(fR);
        if (seen[(size_t)eL * 4 + sL]++ || seen[(size_t)eR * 4 + sR]++)
            throw InputError("IFaceData: an element face appears twice");
        // the flux point with face parameter s_q sits at the left face's
        // tangential grid point a and the right face's point b; b = a or
        // m - 1 - a (the pairing of construct_face_states, base.cpp:287-327)
        const unsigned rev = sigma(fL, oL) != sigma(fR, oR) ? 1u : 0u;
        ft.frec[((size_t)eL * 4 + sL) * 2] = (unsigned)eR;
        ft.frec[((size_t)eL * 4 + sL) * 2 + 1] = (unsigned)fR | (rev << 2) | (1u << 3);
        ft.frec[((size_t)eR * 4 + sR) * 2] = (unsigned)eL;
        ft.frec[((size_t)eR * 4 + sR) * 2 + 1] = (unsigned)fL | (rev << 2);
    }
    for (size_t i = 0; i < seen.size(); ++i)
        if (!seen[i])
            throw UnsupportedError("element " + std::to_string(i / 4) +
                                   " has a face without a neighbour: boundary faces are not on this path "
                                   "(periodic meshes only)");
    return ft;
}

KIC make_ic(const CaseConfig& c) {
    KIC k;
    std::memset(&k, 0, sizeof(k));
    k.ic = (int)c.ic;
    for (size_t i = 0; i < c.ic_data.size() && i < 24; ++i) k.data[i] = c.ic_data[i];
    return k;
}

}  // namespace dg

"""2D (quadrilateral) oracle -- TEST INFRASTRUCTURE.

A numpy restatement of the reference's dim = 2 path on periodic boxes of
affine quadrilaterals, vectorised over elements.  Only tests/ may import
this module; it never backs the product path.

Followed from the reference (paths relative to its root):
  * state U[E][nb][4] = (rho, rho u, rho v, rho E), basis b = j (p+1) + i
    (src/numerics/basis/tools.cpp:150-214, 369-392), Lagrange on
    equidistant nodes or unnormalised Legendre in 1D (the 1D functions of
    the 3D oracle, basis.cpp:15-90)
  * quadrature orders: volume from QuadrilateralQuadrature (segment order
    2p+1, plus 2: p+2 points per direction, quadrilateral.cpp:36-41), faces
    from the segment rule of order 2p+1 (p+1 points, segment.cpp:424-426,
    solver/helpers.cpp:225), mass / IC projection / L2 error from order 2p
    ((2p+2) points, base.cpp:78,129, post/main.cpp:64-67); the point count
    rule is quadrature/tools.cpp:21-25, point q = j m + i
    (quadrilateral.cpp:26-33)
  * the quadrilateral (shape.h:32-40, shape.cpp:28-200): faces 0 (n0 n1),
    1 (n1 n3), 2 (n3 n2), 3 (n2 n0), counter-clockwise; orientation 1
    reverses a face's points; normal N = (dy/ds, -dx/ds) of the left face
    (|N| = half the edge length)
  * residual (base.cpp:401-561, helpers.cpp:185-198, tools.cpp:6-63,
    flux_functors_impl.h:95-102,168-264): weak volume term with the
    contravariant flux F_j (J^-1)_dj |J| w, face fluxes -Fhat on the left
    and +Fhat on the right element, lifted with the face basis values and
    weights
  * numerical fluxes: the 3D oracle's LLF / HLLC / Roe on the embedded
    state (rho, rho u, rho v, 0, rho E) and normal (Nx, Ny, 0) -- with a
    zero third component the 3D formulas are the 2D ones term by term
    (euler/functions.cpp:200-296 is the 2D HLLC)
  * M^-1 from the (2p+2)^2 rule (basis/tools.cpp:85-110), inverted by LU;
    FE / RK4 in the reference's sequence (stepper.cpp:23-142)
  * initial conditions of Physics<2>::call_IC (physics/base/base.cpp:77-
    88,141-178; euler/functions.cpp:94-145): Uniform, Gaussian and
    IsentropicVortex (the Shu vortex, not wrapped periodically)
  * L2 error normalised by the area, per variable (post/main.cpp:30-201)
"""
import numpy as np

import oracle as O

NS2 = 4
FLUX_ID = {"LaxFriedrichs": 0, "HLLC": 1, "Roe": 2}
BASIS_1D = {"LagrangeQuad": "LagrangeHex", "LegendreQuad": "LegendreHex"}


def nq1d(qorder):
    """QuadratureTools::get_number_of_quadrature_points, 1D count."""
    n = 0
    if (qorder + 2) % 2 == 0:
        n = (qorder + 2) // 2
    if (qorder + 1) % 2 == 0:
        n = (qorder + 1) // 2
    return n


def quad_qorder(order):
    return (2 * order + 1) + 2        # quadrilateral.cpp:36-41


def seg_qorder(order):
    return 2 * order + 1              # segment.cpp:424-426


def basis_1d(basis, p, x):
    v = np.zeros((len(x), p + 1))
    g = np.zeros((len(x), p + 1))
    for i, xi in enumerate(x):
        v[i], g[i] = O.basis_1d(BASIS_1D[basis], p, xi)
    return v, g


def basis_2d(basis, p, pts):
    """val [n][nb], grad [n][nb][2] at reference points pts [n][2]."""
    vx, gx = basis_1d(basis, p, pts[:, 0])
    vy, gy = basis_1d(basis, p, pts[:, 1])
    n, n1 = len(pts), p + 1
    val = np.einsum("qj,qi->qji", vy, vx).reshape(n, n1 * n1)
    grad = np.stack([np.einsum("qj,qi->qji", vy, gx).reshape(n, n1 * n1),
                     np.einsum("qj,qi->qji", gy, vx).reshape(n, n1 * n1)], axis=-1)
    return val, grad


def tensor_rule(n):
    x, w = O.gl_rule(n)
    pts = np.array([(x[i], x[j]) for j in range(n) for i in range(n)])
    wts = np.array([w[i] * w[j] for j in range(n) for i in range(n)])
    return pts, wts


# reference corner coordinates of n0..n3 and the nodes of each face
XREF = np.array([[-1., -1.], [1., -1.], [-1., 1.], [1., 1.]])
FACE_NODES = [(0, 1), (1, 3), (3, 2), (2, 0)]


def points_on_face(face, orient, s):
    """get_points_on_face_quadrilateral (shape.cpp:69-122)."""
    a, b = FACE_NODES[face]
    if orient == 1:
        a, b = b, a
    b1 = (np.asarray(s) + 1.) / 2.
    b0 = 1. - b1
    return b0[:, None] * XREF[a] + b1[:, None] * XREF[b]


def embed(U):
    """(rho, m_x, m_y, E) -> (rho, m_x, m_y, 0, E)."""
    U = np.asarray(U, dtype=np.float64)
    out = np.zeros(U.shape[:-1] + (5,))
    out[..., :3] = U[..., :3]
    out[..., 4] = U[..., 3]
    return out


def drop(F5):
    return np.concatenate([F5[..., :3], F5[..., 4:5]], axis=-1)


def flux_interior(U, gamma):
    """[n][4] -> [n][2][4] (euler/functions.cpp:10-62, dim 2)."""
    U5 = np.ascontiguousarray(embed(U).reshape(-1, 5))
    F = np.zeros((len(U5), 3, 5))
    O.lib().orc_flux_interior_n(gamma, len(U5), U5, F.reshape(-1))
    return drop(F[:, :2, :])


def flux_numerical(name, UL, UR, N, gamma):
    UL5 = np.ascontiguousarray(embed(UL).reshape(-1, 5))
    UR5 = np.ascontiguousarray(embed(UR).reshape(-1, 5))
    N3 = np.zeros((len(UL5), 3))
    N3[:, :2] = np.asarray(N).reshape(-1, 2)
    F = np.zeros((len(UL5), 5))
    O.lib().orc_flux_numerical_n(FLUX_ID[name], gamma, len(UL5), UL5, UR5, np.ascontiguousarray(N3), F)
    return drop(F)


def ic_state(ic, data, gamma, R, x, t=0.0):
    """Physics<2>::call_IC / call_exact_solution at points x [n][2]."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    U = np.zeros((n, NS2))
    if ic == "Uniform":
        U[:] = np.asarray(data[:NS2])
    elif ic == "Gaussian":
        s = np.sin(30. * x[:, 0] * x[:, 0] * x[:, 1]) * (x[:, 0] + 1.) * (x[:, 1] + 1.)
        for i in range(NS2):
            U[:, i] = data[i] * s
    elif ic == "IsentropicVortex":
        rhob, ub, vb, pb, vs = data[:5]
        Tb = pb / (rhob * R)
        s = pb / rhob ** gamma
        xr = x[:, 0] - ub * t
        yr = x[:, 1] - vb * t
        r = np.sqrt(xr ** 2 + yr ** 2)
        dU = vs / (2.0 * np.pi) * np.exp(0.5 * (1.0 - r ** 2))
        du = dU * -1.0 * yr
        dv = dU * xr
        dT = -1.0 * (gamma - 1.0) * vs ** 2 / (8.0 * gamma * np.pi ** 2) * np.exp(1.0 - r ** 2)
        u, v, T = ub + du, vb + dv, Tb + dT
        rho = (T / s) ** (1.0 / (gamma - 1.0))
        U[:, 0] = rho
        U[:, 1] = rho * u
        U[:, 2] = rho * v
        U[:, 3] = rho * R / (gamma - 1.0) * T + 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / rho
    else:
        raise ValueError(f"no 2D initial condition {ic}")
    return U


class Box2D:
    """Periodic nx x ny box of affine quadrilaterals, element e = ix + nx iy."""

    def __init__(self, order, basis, flux, nelem, lo, hi, gamma=1.4, R=1.0, ic="IsentropicVortex",
                 ic_data=(1.0, 1.0, 1.0, 1.0, 5.0)):
        self.p, self.basis, self.flux = order, basis, flux
        self.nx, self.ny = nelem
        self.lo, self.hi = np.asarray(lo, float), np.asarray(hi, float)
        self.gamma, self.R, self.ic, self.ic_data = gamma, R, ic, tuple(ic_data)
        self.E = self.nx * self.ny
        self.nb = (order + 1) ** 2
        self.h = (self.hi - self.lo) / np.array([self.nx, self.ny])
        self.jdet = self.h[0] * self.h[1] / 4.
        self.time = 0.0
        p = order
        # volume rule and basis
        self.vpts, self.vwts = tensor_rule(nq1d(quad_qorder(p)))
        self.vval, self.vgrad = basis_2d(basis, p, self.vpts)
        # mass / projection / error rule
        self.mpts, self.mwts = tensor_rule(nq1d(quad_qorder(2 * p)))
        self.mval, _ = basis_2d(basis, p, self.mpts)
        MM = np.einsum("qa,qb,q->ab", self.mval, self.mval, self.mwts * self.jdet)
        self.iMM = np.linalg.inv(MM)
        # faces: segment rule, traces of faces 1 (x+) / 3 (x-) and 2 (y+) / 0 (y-)
        self.fs, self.fw = O.gl_rule(nq1d(seg_qorder(p)))
        self.fval = {}
        for f in range(4):
            for o in range(2):
                self.fval[(f, o)], _ = basis_2d(basis, p, points_on_face(f, o, self.fs))
        # left-element normals (N = (dy/ds, -dx/ds), ds along the left face)
        self.Nx = np.array([self.h[1] / 2., 0.])      # face 1 of the left element
        self.Ny = np.array([0., self.h[0] / 2.])      # face 2 of the left element
        ix, iy = np.meshgrid(np.arange(self.nx), np.arange(self.ny), indexing="xy")
        e = (ix + self.nx * iy).ravel()
        self.east = ((ix + 1) % self.nx + self.nx * iy).ravel()
        self.north = (ix + self.nx * ((iy + 1) % self.ny)).ravel()
        self.eidx = e
        centers = np.stack([self.lo[0] + (ix.ravel() + 0.5) * self.h[0],
                            self.lo[1] + (iy.ravel() + 0.5) * self.h[1]], axis=1)
        self.centers = np.zeros((self.E, 2))
        self.centers[e] = centers
        self.U = np.zeros((self.E, self.nb, NS2))

    def phys(self, pts):
        """Physical points [E][n][2] of reference points pts [n][2]."""
        return self.centers[:, None, :] + 0.5 * self.h[None, None, :] * pts[None, :, :]

    # -- Solver<2> surface --
    def init_state(self):
        x = self.phys(self.mpts).reshape(-1, 2)
        u = ic_state(self.ic, self.ic_data, self.gamma, self.R, x).reshape(self.E, -1, NS2)
        rhs = np.einsum("qb,eqs,q->ebs", self.mval, u, self.mwts * self.jdet)
        self.U = np.einsum("ab,ebs->eas", self.iMM, rhs)
        self.time = 0.0

    def residual(self, U):
        U = np.asarray(U, dtype=np.float64).reshape(self.E, self.nb, NS2)
        g = self.gamma
        # volume term
        Uq = np.einsum("qb,ebs->eqs", self.vval, U)
        F = flux_interior(Uq.reshape(-1, NS2), g).reshape(self.E, -1, 2, NS2)
        G = F * (2. / self.h)[None, None, :, None] * (self.jdet * self.vwts)[None, :, None, None]
        R = np.einsum("qbd,eqds->ebs", self.vgrad, G)
        # faces normal to x: left = e (face 1, orient 0), right = east (face 3, orient 1);
        # faces normal to y: left = e (face 2, orient 0), right = north (face 0, orient 1)
        for (fl, fr, nbr, N) in ((1, 3, self.east, self.Nx), (2, 0, self.north, self.Ny)):
            UL = np.einsum("qb,ebs->eqs", self.fval[(fl, 0)], U)
            UR = np.einsum("qb,ebs->eqs", self.fval[(fr, 1)], U[nbr])
            nq = UL.shape[1]
            Fh = flux_numerical(self.flux, UL.reshape(-1, NS2), UR.reshape(-1, NS2),
                                np.tile(N, (self.E * nq, 1)), g).reshape(self.E, nq, NS2)
            Fw = Fh * self.fw[None, :, None]
            R -= np.einsum("qb,eqs->ebs", self.fval[(fl, 0)], Fw)
            np.add.at(R, nbr, np.einsum("qb,eqs->ebs", self.fval[(fr, 1)], Fw))
        return R

    def inv_mass(self, dt, R):
        return dt * np.einsum("ab,ebs->eas", self.iMM, np.asarray(R).reshape(self.E, self.nb, NS2))

    def step(self, stepper, dt, nsteps):
        for _ in range(nsteps):
            if stepper == "FE":
                self.U = self.U + self.inv_mass(dt, self.residual(self.U))
                self.time += dt
            else:
                U0 = self.U.copy()
                d1 = self.inv_mass(dt, self.residual(U0))
                U = U0 + 0.5 * d1
                self.time += dt / 2.
                d2 = self.inv_mass(dt, self.residual(U))
                U = U0 + 0.5 * d2
                d3 = self.inv_mass(dt, self.residual(U))
                U = U0 + 1.0 * d3
                self.time += dt / 2.
                d4 = self.inv_mass(dt, self.residual(U))
                d4 = d4 + 2.0 * d3
                d4 = d4 + 2.0 * d2
                d4 = d4 + 1.0 * d1
                self.U = U0 + (1.0 / 6.0) * d4

    def l2_error(self, t):
        x = self.phys(self.mpts).reshape(-1, 2)
        ue = ic_state(self.ic, self.ic_data, self.gamma, self.R, x, t).reshape(self.E, -1, NS2)
        uh = np.einsum("qb,ebs->eqs", self.mval, self.U)
        w = self.mwts * self.jdet
        num = np.einsum("eqs,q->s", (uh - ue) ** 2, w)
        return np.sqrt(num / (w.sum() * self.E))

    def integrals(self):
        uh = np.einsum("qb,ebs->eqs", self.mval, self.U)
        return np.einsum("eqs,q->s", uh, self.mwts * self.jdet)


# ---------------------------------------------------------------------------
# Quadrilateral meshes given as the reference's mesh arrays (Mesh::read_mesh,
# src/mesh/mesh.cpp:79-175): NodeCoords [n][2], Elem2Nodes [E][4] (node
# 2j + i of the element's frame), IFaceData [F][6] = (elemL, faceL, orientL,
# elemR, faceR, orientR).  Bilinear elements: per-point metric
# (helpers.cpp:4-51, basis/tools.cpp:42-66), dense per-element mass
# matrices (helpers.cpp:53-100), face points from the orientations
# (base.cpp:287-327), normals of the left face N = (dy/ds, -dx/ds)
# (shape.cpp:157-200; straight edges, so constant along the face).
# ---------------------------------------------------------------------------

def bilinear(X, pts):
    """Physical points, Jacobian columns a_xi, a_eta [n][2] and |J| at
    reference points pts [n][2] of an element with nodes X [4][2]."""
    xi, eta = pts[:, 0], pts[:, 1]
    x = np.zeros((len(pts), 2))
    ax = np.zeros((len(pts), 2))
    ae = np.zeros((len(pts), 2))
    for n in range(4):
        sx, se = 2 * (n & 1) - 1, 2 * (n >> 1) - 1
        w = (1 + sx * xi) * (1 + se * eta) / 4.
        dx = sx * (1 + se * eta) / 4.
        de = se * (1 + sx * xi) / 4.
        x += w[:, None] * X[n]
        ax += dx[:, None] * X[n]
        ae += de[:, None] * X[n]
    jd = ax[:, 0] * ae[:, 1] - ax[:, 1] * ae[:, 0]
    return x, ax, ae, jd


class Mesh2D(Box2D):
    """The same solver on a quadrilateral mesh; lo/hi give the periodic cell
    of the exact solutions only."""

    def __init__(self, order, basis, flux, mesh, lo, hi, gamma=1.4, R=1.0, ic="IsentropicVortex",
                 ic_data=(1.0, 1.0, 1.0, 1.0, 5.0)):
        xyz = np.asarray(mesh[0], dtype=np.float64)
        e2n = np.asarray(mesh[1], dtype=np.int64)
        self.ifaces = np.asarray(mesh[2], dtype=np.int64)
        self.p, self.basis, self.flux = order, basis, flux
        self.lo, self.hi = np.asarray(lo, float), np.asarray(hi, float)
        self.gamma, self.R, self.ic, self.ic_data = gamma, R, ic, tuple(ic_data)
        self.E = len(e2n)
        self.nb = (order + 1) ** 2
        self.time = 0.0
        p = order
        self.X = xyz[e2n]                                   # [E][4][2]
        self.vpts, self.vwts = tensor_rule(nq1d(quad_qorder(p)))
        self.vval, self.vgrad = basis_2d(basis, p, self.vpts)
        self.mpts, self.mwts = tensor_rule(nq1d(quad_qorder(2 * p)))
        self.mval, _ = basis_2d(basis, p, self.mpts)
        # volume metric: cofactor rows C_xi = (a_eta.y, -a_eta.x),
        # C_eta = (-a_xi.y, a_xi.x), i.e. |J| (J^-1) rows
        self.cof = np.zeros((self.E, len(self.vpts), 2, 2))
        self.mx = np.zeros((self.E, len(self.mpts), 2))
        self.mjd = np.zeros((self.E, len(self.mpts)))
        self.iMMe = np.zeros((self.E, self.nb, self.nb))
        for e in range(self.E):
            _, ax, ae, jd = bilinear(self.X[e], self.vpts)
            if np.any(jd <= 0):
                raise ValueError(f"element {e} is folded")
            self.cof[e, :, 0, 0], self.cof[e, :, 0, 1] = ae[:, 1], -ae[:, 0]
            self.cof[e, :, 1, 0], self.cof[e, :, 1, 1] = -ax[:, 1], ax[:, 0]
            x, _, _, jdm = bilinear(self.X[e], self.mpts)
            self.mx[e], self.mjd[e] = x, jdm
            MM = np.einsum("qa,qb,q->ab", self.mval, self.mval, self.mwts * jdm)
            self.iMMe[e] = np.linalg.inv(MM)
        self.fs, self.fw = O.gl_rule(nq1d(seg_qorder(p)))
        self.U = np.zeros((self.E, self.nb, NS2))

    def init_state(self):
        u = ic_state(self.ic, self.ic_data, self.gamma, self.R, self.mx.reshape(-1, 2)).reshape(self.E, -1, NS2)
        rhs = np.einsum("qb,eqs,eq->ebs", self.mval, u, self.mwts[None, :] * self.mjd)
        self.U = np.einsum("eab,ebs->eas", self.iMMe, rhs)
        self.time = 0.0

    def residual(self, U):
        U = np.asarray(U, dtype=np.float64).reshape(self.E, self.nb, NS2)
        g = self.gamma
        Uq = np.einsum("qb,ebs->eqs", self.vval, U)
        F = flux_interior(Uq.reshape(-1, NS2), g).reshape(self.E, -1, 2, NS2)     # [e][q][j][s]
        G = np.einsum("eqdj,eqjs->eqds", self.cof, F) * self.vwts[None, :, None, None]
        R = np.einsum("qbd,eqds->ebs", self.vgrad, G)
        for eL, fL, oL, eR, fR, oR in self.ifaces:
            vL, _ = basis_2d(self.basis, self.p, points_on_face(fL, oL, self.fs))
            vR, _ = basis_2d(self.basis, self.p, points_on_face(fR, oR, self.fs))
            UL, UR = vL @ U[eL], vR @ U[eR]
            a, b = FACE_NODES[fL]
            d = (self.X[eL, b] - self.X[eL, a]) / 2.
            N = np.tile([d[1], -d[0]], (len(self.fs), 1))
            Fw = flux_numerical(self.flux, UL, UR, N, g) * self.fw[:, None]
            R[eL] -= vL.T @ Fw
            R[eR] += vR.T @ Fw
        return R

    def inv_mass(self, dt, R):
        return dt * np.einsum("eab,ebs->eas", self.iMMe, np.asarray(R).reshape(self.E, self.nb, NS2))

    def l2_error(self, t):
        ue = ic_state(self.ic, self.ic_data, self.gamma, self.R, self.mx.reshape(-1, 2), t).reshape(self.E, -1, NS2)
        uh = np.einsum("qb,ebs->eqs", self.mval, self.U)
        w = self.mwts[None, :] * self.mjd
        return np.sqrt(np.einsum("eqs,eq->s", (uh - ue) ** 2, w) / w.sum())

    def integrals(self):
        uh = np.einsum("qb,ebs->eqs", self.mval, self.U)
        return np.einsum("eqs,eq->s", uh, self.mwts[None, :] * self.mjd)

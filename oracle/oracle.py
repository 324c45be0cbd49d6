"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, the smoke check and the benchmark's CPU-baseline leg may import
this module.  It never backs the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BASIS = {"LagrangeHex": 0, "LegendreHex": 1}
FLUX = {"LaxFriedrichs": 0, "HLLC": 1, "Roe": 2}
IC = {"Uniform": 0, "Sphere": 1, "IsentropicVortex": 2, "ExtrudedIsentropicVortex": 3,
      "GaussianBump": 4, "TaylorGreen": 5}
STEPPER = {"FE": 0, "RK4": 1}


class Params(C.Structure):
    _fields_ = [("order", C.c_int), ("basis", C.c_int), ("flux", C.c_int),
                ("nelem", C.c_int * 3), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
                ("gamma", C.c_double), ("R", C.c_double), ("ic", C.c_int),
                ("n_ic_data", C.c_int), ("ic_data", C.c_double * 10), ("warp", C.c_double),
                ("exact_normals", C.c_int)]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")

_lib = None


def build():
    """Compile liboracle.so with its Makefile (gcc)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_nq1d_from_qorder": (C.c_int, [C.c_int]),
            "orc_hex_qorder": (C.c_int, [C.c_int]),
            "orc_quad_face_qorder": (C.c_int, [C.c_int]),
            "orc_gl_rule": (None, [C.c_int, _dp, _dp]),
            "orc_equidistant_nodes": (None, [C.c_double, C.c_double, C.c_int, _dp]),
            "orc_basis_1d": (None, [C.c_int, C.c_int, C.c_double, _dp, _dp]),
            "orc_basis_3d": (None, [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]),
            "orc_face_pts_order": (None, [C.c_int, C.c_int, _ip]),
            "orc_points_on_face": (None, [C.c_int, C.c_int, C.c_int, _dp, _dp]),
            "orc_local_nodes_on_face": (None, [C.c_int, C.c_int, _ip]),
            "orc_pressure": (C.c_double, [C.c_double, _dp]),
            "orc_maxwavespeed": (C.c_double, [C.c_double, _dp]),
            "orc_flux_interior": (None, [C.c_double, _dp, _dp]),
            "orc_flux_numerical": (None, [C.c_int, C.c_double, _dp, _dp, _dp, _dp]),
            "orc_flux_interior_n": (None, [C.c_double, C.c_int, _dp, _dp]),
            "orc_flux_numerical_n": (None, [C.c_int, C.c_double, C.c_int, _dp, _dp, _dp, _dp]),
            "orc_ic_eval": (None, [C.POINTER(Params), _dp, C.c_double, _dp]),
            "orc_create": (C.c_void_p, [C.POINTER(Params)]),
            "orc_create_mesh": (C.c_void_p, [C.POINTER(Params), C.c_long, _dp, C.c_long,
                                             np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                             C.c_long, _ip]),
            "orc_destroy": (None, [C.c_void_p]),
            "orc_num_elems": (C.c_long, [C.c_void_p]),
            "orc_num_basis": (C.c_int, [C.c_void_p]),
            "orc_num_ifaces": (C.c_long, [C.c_void_p]),
            "orc_get_face_table": (None, [C.c_void_p, _up]),
            "orc_get_normals": (None, [C.c_void_p, _dp]),
            "orc_get_imm": (None, [C.c_void_p, C.c_long, _dp]),
            "orc_init_state": (None, [C.c_void_p]),
            "orc_set_state": (None, [C.c_void_p, _dp]),
            "orc_get_state": (None, [C.c_void_p, _dp]),
            "orc_residual": (None, [C.c_void_p, _dp, _dp]),
            "orc_inv_mass": (None, [C.c_void_p, C.c_double, _dp, _dp]),
            "orc_step": (None, [C.c_void_p, C.c_int, C.c_double, C.c_int]),
            "orc_time": (C.c_double, [C.c_void_p]),
            "orc_l2_error": (None, [C.c_void_p, C.c_double, _dp]),
            "orc_max_wavespeed_state": (C.c_double, [C.c_void_p]),
            "orc_total_volume": (C.c_double, [C.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def make_params(order=2, basis="LegendreHex", flux="LaxFriedrichs", nelem=(2, 2, 2),
                lo=(0., 0., 0.), hi=(1., 1., 1.), gamma=1.4, R=1.0,
                ic="Uniform", ic_data=(1.0, 0.1, 0.2, 0.3, 2.5), warp=0.0, exact_normals=False):
    p = Params()
    p.warp = float(warp)
    p.exact_normals = 1 if exact_normals else 0
    p.order = order
    p.basis = BASIS[basis]
    p.flux = FLUX[flux]
    for d in range(3):
        p.nelem[d] = int(nelem[d])
        p.lo[d] = float(lo[d])
        p.hi[d] = float(hi[d])
    p.gamma = gamma
    p.R = R
    p.ic = IC[ic]
    p.n_ic_data = len(ic_data)
    for i, v in enumerate(ic_data):
        p.ic_data[i] = float(v)
    return p


# ---------------- thin numpy wrappers ----------------

def gl_rule(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gl_rule(n, x, w)
    return x, w


def basis_3d(basis, p, pts):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    nb = (p + 1) ** 3
    val = np.zeros((len(pts), nb))
    grad = np.zeros((len(pts), nb, 3))
    lib().orc_basis_3d(BASIS[basis], p, len(pts), pts, val, grad)
    return val, grad


def basis_1d(basis, p, x):
    val = np.zeros(p + 1)
    grad = np.zeros(p + 1)
    lib().orc_basis_1d(BASIS[basis], p, float(x), val, grad)
    return val, grad


def face_pts_order(orient, npts):
    o = np.zeros(npts, dtype=np.int32)
    lib().orc_face_pts_order(orient, npts, o)
    return o


def points_on_face(face, orient, face_pts):
    fp = np.ascontiguousarray(face_pts, dtype=np.float64).reshape(-1, 2)
    out = np.zeros((len(fp), 3))
    lib().orc_points_on_face(face, orient, len(fp), fp, out)
    return out


def flux_interior(U, gamma=1.4):
    F = np.zeros(15)
    lib().orc_flux_interior(gamma, np.ascontiguousarray(U, dtype=np.float64), F)
    return F.reshape(3, 5)


def flux_numerical(name, UL, UR, N, gamma=1.4):
    F = np.zeros(5)
    lib().orc_flux_numerical(FLUX[name], gamma, np.ascontiguousarray(UL, dtype=np.float64),
                             np.ascontiguousarray(UR, dtype=np.float64),
                             np.ascontiguousarray(N, dtype=np.float64), F)
    return F


def pressure(U, gamma=1.4):
    return lib().orc_pressure(gamma, np.ascontiguousarray(U, dtype=np.float64))


def maxwavespeed(U, gamma=1.4):
    return lib().orc_maxwavespeed(gamma, np.ascontiguousarray(U, dtype=np.float64))


def ic_eval(params, x, t=0.0):
    U = np.zeros(5)
    lib().orc_ic_eval(C.byref(params), np.ascontiguousarray(x, dtype=np.float64), t, U)
    return U


class Solver:
    """Oracle solver on a periodic box (reference semantics)."""

    def __init__(self, params, mesh=None):
        """mesh: (NodeCoords [n][3], Elem2Nodes [E][8], IFaceData [F][6]) in
        the reference's mesh-file layout; None: the periodic box of params."""
        self.params = params
        if mesh is None:
            self.h = lib().orc_create(C.byref(params))
        else:
            xyz = np.ascontiguousarray(mesh[0], dtype=np.float64)
            e2n = np.ascontiguousarray(mesh[1], dtype=np.int64)
            ifd = np.ascontiguousarray(mesh[2], dtype=np.int32)
            self.h = lib().orc_create_mesh(C.byref(params), len(xyz), xyz.ravel(), len(e2n), e2n.ravel(),
                                           len(ifd), ifd.ravel())
        self.E = lib().orc_num_elems(self.h)
        self.nb = lib().orc_num_basis(self.h)
        self.F = lib().orc_num_ifaces(self.h)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self):
        return (self.E, self.nb, 5)

    def init_state(self):
        lib().orc_init_state(self.h)

    def set_state(self, U):
        lib().orc_set_state(self.h, np.ascontiguousarray(U, dtype=np.float64).reshape(self.shape()))

    def get_state(self):
        U = np.zeros(self.shape())
        lib().orc_get_state(self.h, U)
        return U

    def residual(self, U):
        R = np.zeros(self.shape())
        lib().orc_residual(self.h, np.ascontiguousarray(U, dtype=np.float64).reshape(self.shape()), R)
        return R

    def inv_mass(self, dt, R):
        dU = np.zeros(self.shape())
        lib().orc_inv_mass(self.h, dt, np.ascontiguousarray(R, dtype=np.float64), dU)
        return dU

    def step(self, stepper, dt, nsteps):
        lib().orc_step(self.h, STEPPER[stepper], dt, nsteps)

    @property
    def time(self):
        return lib().orc_time(self.h)

    def l2_error(self, t):
        e = np.zeros(5)
        lib().orc_l2_error(self.h, t, e)
        return e

    def max_wavespeed(self):
        return lib().orc_max_wavespeed_state(self.h)

    def total_volume(self):
        return lib().orc_total_volume(self.h)

    def face_table(self):
        t = np.zeros((self.F, 8), dtype=np.uint32)
        lib().orc_get_face_table(self.h, t)
        return t

    def normals(self, nqf):
        n = np.zeros((self.F, nqf, 3))
        lib().orc_get_normals(self.h, n)
        return n

    def imm(self, e):
        m = np.zeros((self.nb, self.nb))
        lib().orc_get_imm(self.h, e, m)
        return m

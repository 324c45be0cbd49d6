"""Python binding of the MI355X DG residual + RK path (libdg_amd.so).

Mirrors the reference's solver surface (src/solver/base.h:16-90,
src/numerics/timestepping/stepper.h:14-90, src/post/main.cpp:30-201):

    s = Solver(toml_text)          # Solver<3>(input, mesh, ...) + precompute
    s.init_state_from_fcn()        # L2 projection of the initial condition
    s.get_residual(U_ptr, R_ptr)   # R = volume - face (device buffers)
    s.take_time_step()             # one FE / RK4 step
    s.solve()                      # (int)(end / dt) steps
    s.get_error()                  # volume-normalised L2 error per variable

The compute path is the HIP library only: importing fails when it is not
built, and creating a solver fails without a GPU.  There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
# DG_AMD_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("DG_AMD_LIB") or os.path.join(PKG_ROOT, "lib", "libdg_amd.so")

DG_LAYOUT_RIGHT = 0
DG_LAYOUT_LEFT = 1
STATUS = {0: "DG_OK", 1: "DG_ERR_INPUT", 2: "DG_ERR_DEVICE", 3: "DG_ERR_ARGUMENT",
          4: "DG_ERR_UNSUPPORTED", 5: "DG_ERR_NUMERICS"}
FLUX_NAMES = ["LaxFriedrichs", "HLLC", "Roe"]
BASIS_NAMES = ["LagrangeHex", "LegendreHex"]
STEPPER_NAMES = ["FE", "RK4"]

# every symbol include/dg_amd.h declares
EXPORTS = ["dg_create", "dg_create_mesh", "dg_destroy", "dg_last_error", "dg_get_info", "dg_get_mesh", "dg_init_state",
           "dg_set_state", "dg_get_state", "dg_state_device_ptr", "dg_residual",
           "dg_apply_inv_mass", "dg_step", "dg_solve", "dg_l2_error", "dg_integrals",
           "dg_max_wavespeed", "dg_set_time", "dg_set_stream", "dg_batch_create", "dg_batch_step",
           "dg_batch_destroy", "dg_batch_last_error", "dg_batch_tuning", "dg_batch_tuning_tiles", "dg_mlp_create", "dg_mlp_destroy", "dg_mlp_last_error",
           "dg_mlp_num_params", "dg_mlp_get_params", "dg_mlp_set_params", "dg_mlp_set_scaling", "dg_mlp_forward",
           "dg_mlp_gradient", "dg_mlp_train", "dg_closure_features", "dg_closure_target", "dg_closure_target_at",
           "dg_closure_step", "dg_closure_step_noalloc"]


class DGError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Info(C.Structure):
    _fields_ = [("num_elems", C.c_long), ("order", C.c_int), ("num_basis", C.c_int),
                ("num_state_vars", C.c_int), ("nelem", C.c_int * 3),
                ("lower", C.c_double * 3), ("upper", C.c_double * 3),
                ("gamma", C.c_double), ("gas_constant", C.c_double),
                ("timestep", C.c_double), ("end_time", C.c_double), ("time", C.c_double),
                ("num_iterations", C.c_int), ("stepper", C.c_int), ("flux", C.c_int),
                ("basis", C.c_int), ("ic", C.c_int), ("dim", C.c_int), ("precision", C.c_int)]


_lib = None


def lib():
    """Load libdg_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                              "(make -C dg-mpi_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        sig = {
            "dg_create": [C.c_char_p, C.POINTER(vp)],
            "dg_create_mesh": [C.c_char_p, C.c_long, dp, C.c_long, C.POINTER(C.c_long), C.c_long,
                               C.POINTER(C.c_int), C.POINTER(vp)],
            "dg_destroy": [vp],
            "dg_get_info": [vp, C.POINTER(Info)],
            "dg_get_mesh": [vp, C.POINTER(C.c_long), dp, C.POINTER(C.c_long)],
            "dg_init_state": [vp],
            "dg_set_state": [vp, dp, C.c_int],
            "dg_get_state": [vp, dp, C.c_int],
            "dg_state_device_ptr": [vp, C.POINTER(dp)],
            "dg_residual": [vp, vp, vp, vp],
            "dg_apply_inv_mass": [vp, C.c_double, vp, vp, vp],
            "dg_step": [vp, C.c_double, C.c_int],
            "dg_solve": [vp],
            "dg_l2_error": [vp, C.c_double, dp],
            "dg_integrals": [vp, dp],
            "dg_max_wavespeed": [vp, dp],
            "dg_set_time": [vp, C.c_double],
            "dg_set_stream": [vp, vp],
            "dg_batch_create": [C.c_int, C.POINTER(vp), C.POINTER(vp)],
            "dg_batch_step": [vp, C.c_int],
            "dg_batch_destroy": [vp],
            "dg_batch_tuning": [vp, C.POINTER(C.c_int), vp, vp, vp, vp],
            "dg_batch_tuning_tiles": [vp, C.POINTER(C.c_int), vp, vp, vp, vp, vp],
            "dg_mlp_create": [C.c_ulonglong, C.POINTER(vp)],
            "dg_mlp_destroy": [vp],
            "dg_mlp_num_params": [C.POINTER(C.c_long)],
            "dg_mlp_get_params": [vp, vp],
            "dg_mlp_set_params": [vp, vp],
            "dg_mlp_set_scaling": [vp, vp, vp, vp],
            "dg_mlp_forward": [vp, vp, vp, C.c_long],
            "dg_mlp_gradient": [vp, vp, vp, C.c_long, vp, dp],
            "dg_mlp_train": [vp, vp, vp, C.c_long, vp, C.c_long, C.c_int, C.c_float, C.c_float, dp],
            "dg_closure_features": [vp, vp, vp],
            "dg_closure_target": [vp, vp, vp, vp, vp],
            "dg_closure_target_at": [vp, vp, vp, vp, C.c_double, vp],
            "dg_closure_step": [vp, vp, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_long)],
            "dg_closure_step_noalloc": [vp, vp, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_long)],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.dg_last_error.argtypes = [vp]
        L.dg_last_error.restype = C.c_char_p
        L.dg_batch_last_error.argtypes = [vp]
        L.dg_batch_last_error.restype = C.c_char_p
        L.dg_mlp_last_error.argtypes = [vp]
        L.dg_mlp_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Solver:
    """One DG case on the current HIP device."""

    def __init__(self, toml_text, mesh=None):
        """mesh: None for the periodic box of the TOML case, else (node
        coordinates [n][3], element nodes [E][8], interior faces [F][6])
        in the layout of the reference's mesh file; for Physics.dim = 2,
        coordinates [n][2] and element nodes [E][4]."""
        h = C.c_void_p()
        if mesh is None:
            rc = lib().dg_create(toml_text.encode(), C.byref(h))
        else:
            xyz = np.ascontiguousarray(mesh[0], dtype=np.float64)
            e2n = np.ascontiguousarray(mesh[1], dtype=np.int64)
            ifc = np.ascontiguousarray(mesh[2], dtype=np.int32)
            dim = xyz.shape[1] if xyz.ndim == 2 else 0
            if dim not in (2, 3) or e2n.ndim != 2 or e2n.shape[1] != 2 ** dim or ifc.ndim != 2 or \
                    ifc.shape[1] != 6:
                raise ValueError("mesh arrays must be [n][3], [E][8] and [F][6] (2D: [n][2], [E][4], [F][6])")
            rc = lib().dg_create_mesh(toml_text.encode(), xyz.shape[0], _ptr(xyz), e2n.shape[0],
                                      e2n.ctypes.data_as(C.POINTER(C.c_long)), ifc.shape[0],
                                      ifc.ctypes.data_as(C.POINTER(C.c_int)), C.byref(h))
        if rc != 0:
            raise DGError(rc, lib().dg_last_error(None).decode())
        self._h = h
        self.info = self._info()

    @classmethod
    def _adopt(cls, h):
        """Wrap a handle created elsewhere (libdg_amd_h5)."""
        self = cls.__new__(cls)
        self._h = h
        self.info = self._info()
        return self

    # -- plumbing --
    def _check(self, rc):
        if rc != 0:
            raise DGError(rc, lib().dg_last_error(self._h).decode())

    def _info(self):
        i = Info()
        self._check(lib().dg_get_info(self._h, C.byref(i)))
        return i

    def close(self):
        if getattr(self, "_h", None):
            lib().dg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def shape(self):
        i = self.info
        return (i.num_elems, i.num_basis, i.num_state_vars)

    @property
    def time(self):
        return self._info().time

    @property
    def timestep(self):
        return self._info().timestep

    @property
    def num_iterations(self):
        return self._info().num_iterations

    def mesh(self):
        """(node coordinates [n][dim], element nodes [E][2^dim]) of the case."""
        n = C.c_long()
        self._check(lib().dg_get_mesh(self._h, C.byref(n), None, None))
        dim = self.info.dim
        xyz = np.zeros((n.value, dim))
        e2n = np.zeros((self.info.num_elems, 2 ** dim), dtype=np.int64)
        self._check(lib().dg_get_mesh(self._h, C.byref(n), _ptr(xyz), e2n.ctypes.data_as(C.POINTER(C.c_long))))
        return xyz, e2n

    # -- reference surface --
    def init_state_from_fcn(self):
        self._check(lib().dg_init_state(self._h))
        self.info = self._info()

    def get_state(self, layout=DG_LAYOUT_RIGHT):
        i = self.info
        U = np.zeros(i.num_elems * i.num_basis * i.num_state_vars)
        self._check(lib().dg_get_state(self._h, _ptr(U), layout))
        return U.reshape(self.shape) if layout == DG_LAYOUT_RIGHT else U.reshape(
            (i.num_state_vars, i.num_basis, i.num_elems))

    def set_state(self, U, layout=DG_LAYOUT_RIGHT):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.size != int(np.prod(self.shape)):
            raise ValueError("state has the wrong size")
        self._check(lib().dg_set_state(self._h, _ptr(U), layout))

    def state_device_ptr(self):
        p = C.POINTER(C.c_double)()
        self._check(lib().dg_state_device_ptr(self._h, C.byref(p)))
        return C.cast(p, C.c_void_p).value

    def get_residual(self, U_dev, R_dev, stream=None):
        """R = volume - face at U (device pointers, [E][nb][5] FP64)."""
        self._check(lib().dg_residual(self._h, C.c_void_p(U_dev), C.c_void_p(R_dev),
                                      C.c_void_p(stream) if stream else None))

    def mult_inv_mass_matrix(self, dt, R_dev, dU_dev, stream=None):
        self._check(lib().dg_apply_inv_mass(self._h, dt, C.c_void_p(R_dev), C.c_void_p(dU_dev),
                                            C.c_void_p(stream) if stream else None))

    def take_time_step(self, dt=0.0, nsteps=1):
        self._check(lib().dg_step(self._h, dt, nsteps))

    def solve(self):
        self._check(lib().dg_solve(self._h))

    def get_error(self, t=None):
        e = np.zeros(5)
        self._check(lib().dg_l2_error(self._h, self.time if t is None else t, _ptr(e)))
        return e

    def integrals(self):
        o = np.zeros(6)
        self._check(lib().dg_integrals(self._h, _ptr(o)))
        return o

    def max_wavespeed(self):
        a = np.zeros(1)
        self._check(lib().dg_max_wavespeed(self._h, _ptr(a)))
        return float(a[0])

    def set_time(self, t):
        self._check(lib().dg_set_time(self._h, t))

    def set_stream(self, stream):
        """Launch on a caller-owned HIP stream handle (int), None = own."""
        self._check(lib().dg_set_stream(self._h, C.c_void_p(stream) if stream else None))

    # -- data.h5 files (libdg_amd_h5.so) --
    def write(self, path="data.h5"):
        """Write the state and time in the reference's data.h5 layout."""
        from . import h5
        h5.write_solution(self, path)

    def read_in_coefficients(self, path="data.h5", set_time=True):
        """Load a data.h5 file (Solver::read_in_coefficients)."""
        from . import h5
        h5.read_solution(self, path, set_time)


class Batch:
    """Independent box cases advanced together (dg_batch_*): one launch per
    (order, flux, precision) group and RK stage, steps replayed from a
    captured HIP graph.  Borrows the Solver handles; close it first."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s._h.value for s in self.solvers])
        h = C.c_void_p()
        rc = lib().dg_batch_create(len(self.solvers), arr, C.byref(h))
        if rc != 0:
            raise DGError(rc, lib().dg_batch_last_error(None).decode())
        self._h = h

    def step(self, nsteps=1):
        rc = lib().dg_batch_step(self._h, nsteps)
        if rc != 0:
            raise DGError(rc, lib().dg_batch_last_error(self._h).decode())
        for s in self.solvers:
            s.info = s._info()

    def tuning(self):
        """per group: (order, team size in use, ms per step with the default
        and with the alternate team; 0 before tuning)"""
        n = C.c_int()
        rc = lib().dg_batch_tuning(self._h, C.byref(n), None, None, None, None)
        if rc != 0:
            raise DGError(rc, lib().dg_batch_last_error(self._h).decode())
        o, t = (C.c_int * n.value)(), (C.c_int * n.value)()
        a, b = (C.c_double * n.value)(), (C.c_double * n.value)()
        lib().dg_batch_tuning(self._h, C.byref(n), o, t, a, b)
        return [(o[i], t[i], a[i], b[i]) for i in range(n.value)]

    def tiles(self):
        """per group: (brick cap in use, cases, elements, [ms per step of
        each candidate: team variant + 2 cap index]; zeros before tuning)"""
        n, nc = C.c_int(), C.c_int()
        rc = lib().dg_batch_tuning_tiles(self._h, C.byref(n), C.byref(nc), None, None, None, None)
        if rc != 0:
            raise DGError(rc, lib().dg_batch_last_error(self._h).decode())
        cap, nca = (C.c_int * n.value)(), (C.c_int * n.value)()
        el, ms = (C.c_long * n.value)(), (C.c_double * (n.value * nc.value))()
        lib().dg_batch_tuning_tiles(self._h, C.byref(n), C.byref(nc), cap, nca, el, ms)
        return [(cap[i], nca[i], el[i], list(ms[i * nc.value:(i + 1) * nc.value])) for i in range(n.value)]

    def close(self):
        if self._h:
            lib().dg_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def case_toml(order=3, basis="LegendreHex", flux="Roe", nelem=(8, 8, 8), lower=(-5., -5., -5.),
              upper=(5., 5., 5.), stepper="RK4", timestep=None, cfl=None, end=1.0, gamma=1.4,
              gas_constant=1.0, ic="ExtrudedIsentropicVortex", ic_data=(1.0, 1.0, 1.0, 1.0, 5.0),
              warp=0.0, normals="reference", box=True, precision=None, compensated=None, subtract_mean=None):
    """TOML text of a periodic-box case (keys of the reference input decks;
    warp > 0 gives the non-affine box; two entries in nelem, lower and upper
    give the 2D case, Physics.dim = 2).  box=False: the case of a mesh
    passed to Solver(..., mesh=...), whose [Mesh] holds only the normals
    placement and the periodic cell (lower, upper) of the exact solutions.
    precision="FP32": Numerics.precision of the FP32 form (3D box).
    compensated=True/False: Stepper.compensated (default: on for the
    reduced-precision forms).  subtract_mean=True/False:
    Numerics.subtract_mean (default: on for the reduced-precision forms)."""
    st = f"timestep = {timestep!r}" if timestep is not None else f"cfl = {cfl!r}"
    if compensated is not None:
        st += f"\ncompensated = {'true' if compensated else 'false'}"
    prec = f'\nprecision = "{precision}"' if precision else ""
    if subtract_mean is not None:
        prec += f"\nsubtract_mean = {'true' if subtract_mean else 'false'}"
    arr = lambda v: "[" + ", ".join(repr(float(x)) for x in v) + "]"
    # 2: quadrilaterals (basis *Quad, 4 state variables)
    dim = len(nelem) if box else len(lower)
    mesh = (f"""type = "PeriodicBox"
warp = {float(warp)!r}
nelem = [{", ".join(str(int(n)) for n in nelem)}]
""" if box else "")
    return f"""[Mesh]
{mesh}normals = "{normals}"
lower = {arr(lower)}
upper = {arr(upper)}

[Numerics]
order = {order}
basis = "{basis}"{prec}

[Stepper]
type = "{stepper}"
{st}
end = {float(end)!r}

[Physics]
name = "Euler"
dim = {dim}
gamma = {gamma!r}
GasConstant = {gas_constant!r}
convective_flux_fcn = "{flux}"

[InitialCondition]
name = "{ic}"
data = {arr(ic_data)}
"""

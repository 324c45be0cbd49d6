"""Binding of libdg_amd_h5.so: solution files in the reference's data.h5
layout (src/io/writer.cpp:10-98, src/io/HDF5Wrapper.h:780-877; read back as
src/solver/base.cpp:233-279 reads them).

    h5.write_box(path, nelem, lower, upper, Uc, time, layout)   # any host array
    hdr = h5.read_header(path)
    Uc  = h5.read_coefficients(path, layout)
    h5.write_solution(solver, path); h5.read_solution(solver, path)
"""
import ctypes as C
import os

import numpy as np

from . import DG_LAYOUT_LEFT, DG_LAYOUT_RIGHT, PKG_ROOT, STATUS, DGError, lib as _core_lib

H5_LIB_PATH = os.path.join(PKG_ROOT, "lib", "libdg_amd_h5.so")

# every symbol include/dg_amd_h5.h declares
EXPORTS = ["dg_h5_write_box", "dg_h5_read_header", "dg_h5_read_coefficients", "dg_write_solution",
           "dg_read_solution", "dg_h5_last_error"]


class Header(C.Structure):
    _fields_ = [("num_elems", C.c_long), ("num_nodes", C.c_long), ("num_ranks", C.c_int),
                ("dim", C.c_int), ("num_nodes_per_elem", C.c_int), ("num_basis", C.c_int),
                ("num_state_vars", C.c_int), ("stored_layout", C.c_int), ("time", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load libdg_amd_h5.so (raises if it was not built: no HDF5 at build time)."""
    global _lib
    if _lib is None:
        if not os.path.exists(H5_LIB_PATH):
            raise ImportError(f"{H5_LIB_PATH} is not built (needs the HDF5 C library, HDF5_DIR)")
        _core_lib()                 # the core library first, so both share one copy
        L = C.CDLL(H5_LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        sig = {
            "dg_h5_write_box": [C.c_char_p, ip, dp, dp, C.c_int, C.c_int, C.c_double, dp, C.c_int],
            "dg_h5_read_header": [C.c_char_p, C.POINTER(Header)],
            "dg_h5_read_coefficients": [C.c_char_p, dp, C.c_long, C.c_int],
            "dg_write_solution": [vp, C.c_char_p],
            "dg_read_solution": [vp, C.c_char_p, C.c_int],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.dg_h5_last_error.argtypes = []
        L.dg_h5_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise DGError(rc, lib().dg_h5_last_error().decode())


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def write_box(path, nelem, lower, upper, Uc, time=0.0, layout=DG_LAYOUT_RIGHT):
    """Write data.h5 for a periodic box; Uc is [E][nb][ns] (layout RIGHT) or
    [ns][nb][E] (layout LEFT)."""
    Uc = np.ascontiguousarray(Uc, dtype=np.float64)
    if Uc.ndim != 3:
        raise ValueError("Uc must be 3D")
    E = int(np.prod(nelem))
    nb = Uc.shape[1]
    ns = Uc.shape[2] if layout == DG_LAYOUT_RIGHT else Uc.shape[0]
    if (Uc.shape[0] if layout == DG_LAYOUT_RIGHT else Uc.shape[2]) != E:
        raise ValueError("Uc does not have prod(nelem) elements")
    ne = (C.c_int * 3)(*[int(x) for x in nelem])
    lo = np.asarray(lower, dtype=np.float64)
    hi = np.asarray(upper, dtype=np.float64)
    _check(lib().dg_h5_write_box(os.fsencode(path), ne, _dptr(lo), _dptr(hi), nb, ns, float(time),
                                 _dptr(Uc), layout))


def read_header(path):
    h = Header()
    _check(lib().dg_h5_read_header(os.fsencode(path), C.byref(h)))
    return h.as_dict()


def read_coefficients(path, layout=DG_LAYOUT_RIGHT):
    h = read_header(path)
    E, nb, ns = h["num_elems"], h["num_basis"], h["num_state_vars"]
    U = np.empty(E * nb * ns)
    _check(lib().dg_h5_read_coefficients(os.fsencode(path), _dptr(U), U.size, layout))
    return U.reshape((E, nb, ns) if layout == DG_LAYOUT_RIGHT else (ns, nb, E))


def write_solution(solver, path):
    """The solver's state and time to `path` (src/exec/main.cpp:106-112)."""
    _check(lib().dg_write_solution(solver._h, os.fsencode(path)))


def read_solution(solver, path, set_time=True):
    """Solver::read_in_coefficients (src/solver/base.cpp:233-279)."""
    _check(lib().dg_read_solution(solver._h, os.fsencode(path), 1 if set_time else 0))
    solver.info = solver._info()


__all__ = ["write_box", "read_header", "read_coefficients", "write_solution", "read_solution",
           "Header", "EXPORTS", "STATUS", "DG_LAYOUT_LEFT", "DG_LAYOUT_RIGHT"]

"""Binding of libdg_amd_h5.so: solution files in the reference's data.h5
layout (src/io/writer.cpp:10-98, src/io/HDF5Wrapper.h:780-877; read back as
src/solver/base.cpp:233-279 reads them).

    h5.write_box(path, nelem, lower, upper, Uc, time, layout)   # any host array
    hdr = h5.read_header(path)
    Uc  = h5.read_coefficients(path, layout)
    h5.write_solution(solver, path); h5.read_solution(solver, path)
    h5.write_mesh(path, coords, elem_nodes, iface_data)          # mesh files
    info = h5.read_mesh_info(path); coords, e2n, ifaces = h5.read_mesh(path)
    solver = h5.solver_from_mesh_file(toml_text, path)
"""
import ctypes as C
import os

import numpy as np

from . import DG_LAYOUT_LEFT, DG_LAYOUT_RIGHT, PKG_ROOT, STATUS, DGError, lib as _core_lib

H5_LIB_PATH = os.path.join(PKG_ROOT, "lib", "libdg_amd_h5.so")

# every symbol include/dg_amd_h5.h declares
EXPORTS = ["dg_h5_write_box", "dg_h5_read_header", "dg_h5_read_coefficients", "dg_h5_read_mesh_info",
           "dg_h5_read_mesh", "dg_h5_write_mesh", "dg_create_from_mesh_file", "dg_write_solution",
           "dg_read_solution", "dg_h5_last_error"]


class Header(C.Structure):
    _fields_ = [("num_elems", C.c_long), ("num_nodes", C.c_long), ("num_ranks", C.c_int),
                ("dim", C.c_int), ("num_nodes_per_elem", C.c_int), ("num_basis", C.c_int),
                ("num_state_vars", C.c_int), ("stored_layout", C.c_int), ("time", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MeshInfo(C.Structure):
    _fields_ = [(k, C.c_long) for k in ("dim", "qorder", "num_elems", "num_nodes", "num_ifaces",
                                        "num_nodes_per_elem", "num_nodes_per_face", "num_bface_groups")]

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
        vp, dp, ip, lp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        sig = {
            "dg_h5_write_box": [C.c_char_p, ip, dp, dp, C.c_int, C.c_int, C.c_double, dp, C.c_int],
            "dg_h5_read_header": [C.c_char_p, C.POINTER(Header)],
            "dg_h5_read_coefficients": [C.c_char_p, dp, C.c_long, C.c_int],
            "dg_h5_read_mesh_info": [C.c_char_p, C.POINTER(MeshInfo)],
            "dg_h5_read_mesh": [C.c_char_p, dp, C.c_long, lp, C.c_long, ip, C.c_long],
            "dg_h5_write_mesh": [C.c_char_p, C.c_int, C.c_long, dp, C.c_long, lp, C.c_long, ip],
            "dg_create_from_mesh_file": [C.c_char_p, C.c_char_p, C.POINTER(vp)],
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


def read_mesh_info(path):
    m = MeshInfo()
    _check(lib().dg_h5_read_mesh_info(os.fsencode(path), C.byref(m)))
    return m.as_dict()


def read_mesh(path):
    """(NodeCoords [n][dim], Elem2Nodes [E][npe], IFaceData [F][6])."""
    m = read_mesh_info(path)
    xyz = np.zeros((m["num_nodes"], m["dim"]))
    e2n = np.zeros((m["num_elems"], m["num_nodes_per_elem"]), dtype=np.int64)
    ifc = np.zeros((m["num_ifaces"], 6), dtype=np.int32)
    _check(lib().dg_h5_read_mesh(os.fsencode(path), _dptr(xyz), xyz.size, e2n.ctypes.data_as(C.POINTER(C.c_long)),
                                 e2n.size, ifc.ctypes.data_as(C.POINTER(C.c_int)), ifc.size))
    return xyz, e2n, ifc


def write_mesh(path, coords, elem_nodes, iface_data):
    """A mesh file in the layout of the reference's reader: trilinear
    hexahedra (coords [n][3], elem_nodes [E][8]) or bilinear quadrilaterals
    (coords [n][2], elem_nodes [E][4]); iface_data [F][6]."""
    xyz = np.ascontiguousarray(coords, dtype=np.float64)
    e2n = np.ascontiguousarray(elem_nodes, dtype=np.int64)
    ifc = np.ascontiguousarray(iface_data, dtype=np.int32)
    dim = xyz.shape[1] if xyz.ndim == 2 else 0
    if dim not in (2, 3) or e2n.ndim != 2 or e2n.shape[1] != 2 ** dim or ifc.ndim != 2 or ifc.shape[1] != 6:
        raise ValueError("mesh arrays must be [n][3], [E][8] and [F][6] (2D: [n][2], [E][4], [F][6])")
    _check(lib().dg_h5_write_mesh(os.fsencode(path), dim, len(xyz), _dptr(xyz), len(e2n),
                                  e2n.ctypes.data_as(C.POINTER(C.c_long)), len(ifc),
                                  ifc.ctypes.data_as(C.POINTER(C.c_int))))


def solver_from_mesh_file(toml_text, path=None):
    """A Solver on the mesh of `path` (None: the case's Mesh.file)."""
    from . import Solver
    h = C.c_void_p()
    _check(lib().dg_create_from_mesh_file(toml_text.encode(), os.fsencode(path) if path else None, C.byref(h)))
    return Solver._adopt(h)


__all__ = ["write_box", "read_header", "read_coefficients", "write_solution", "read_solution",
           "read_mesh_info", "read_mesh", "write_mesh", "solver_from_mesh_file", "Header", "MeshInfo", "EXPORTS", "STATUS", "DG_LAYOUT_LEFT", "DG_LAYOUT_RIGHT"]

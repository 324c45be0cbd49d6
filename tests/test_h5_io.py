"""data.h5 solution files (libdg_amd_h5.so) on the CPU: the layout of the
reference's Writer (src/io/writer.cpp:10-98) and parallel-dataset convention
(src/io/HDF5Wrapper.h:780-877), and the read-back of
Solver::read_in_coefficients (src/solver/base.cpp:233-279), both layouts.

The structure is checked with the HDF5 command-line dump (h5dump) where it is
installed, independently of the library that wrote the file."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

h5 = pytest.importorskip("dg_amd.h5")
if not os.path.exists(h5.H5_LIB_PATH):
    pytest.skip("libdg_amd_h5.so not built (no HDF5 at build time)", allow_module_level=True)

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def _state(E, nb, ns=5, seed=0):
    return np.random.default_rng(seed).standard_normal((E, nb, ns))


def test_exports_match_header():
    hdr = open(os.path.join(ROOT, "include", "dg_amd_h5.h")).read()
    syms = re.findall(r"^\s*(?:int|const char\*)\s+(dg_\w+)\s*\(", hdr, re.M)
    assert set(syms) == set(h5.EXPORTS)
    L = h5.lib()
    for s in syms:
        assert hasattr(L, s)


@pytest.mark.parametrize("layout", [0, 1])
def test_round_trip_both_layouts(tmp_path, layout):
    nelem, lo, hi = (2, 3, 4), (-1.0, 0.0, 2.0), (1.0, 3.0, 6.0)
    E, nb = 24, 27
    U = _state(E, nb)
    f = str(tmp_path / "data.h5")
    Uw = U if layout == 0 else np.ascontiguousarray(U.transpose(2, 1, 0))
    h5.write_box(f, nelem, lo, hi, Uw, time=0.75, layout=layout)
    h = h5.read_header(f)
    assert h == {"num_elems": E, "num_nodes": 3 * 4 * 5, "num_ranks": 1, "dim": 3,
                 "num_nodes_per_elem": 8, "num_basis": nb, "num_state_vars": 5,
                 "stored_layout": 1 if layout == 0 else 0, "time": 0.75}
    # read back in either layout: bit-exact, transposed as the reference does
    R = h5.read_coefficients(f, 0)
    assert np.array_equal(R, U)
    Lft = h5.read_coefficients(f, 1)
    assert np.array_equal(Lft, U.transpose(2, 1, 0))


def _dump(path, name):
    out = subprocess.run([H5DUMP, "-d", name, "-y", "-w", "0", path], capture_output=True, text=True,
                         check=True).stdout
    body = out[out.index("DATA {") + 6:]
    body = body[:body.index("}")]
    return np.array([float(x) for x in body.replace(",", " ").split()])


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_structure_and_mesh_arrays(tmp_path):
    """Dataset names, element types, the _offsets/_dims companions and the
    box mesh (hex node n = 4k + 2j + i, element e = ix + nx (iy + ny iz))."""
    nelem, lo, hi = (2, 3, 4), (-1.0, 0.0, 2.0), (1.0, 3.0, 6.0)
    E = 24
    f = str(tmp_path / "data.h5")
    h5.write_box(f, nelem, lo, hi, _state(E, 8), time=0.0)
    head = subprocess.run([H5DUMP, "-H", f], capture_output=True, text=True, check=True).stdout

    def dtype_of(kind, name):
        m = re.search(kind + r' "' + re.escape(name) + r'" \{\s*DATATYPE\s+(\S+)\s*DATASPACE\s+([^\n]+)', head)
        assert m, name
        return m.group(1), m.group(2).strip()

    for a in ["Number of Ranks", "Number of Dimensions", "Number of Elements", "Number of Nodes",
              "Number of Nodes Per Element", "Number of Basis Functions", "Number of State Variables"]:
        assert dtype_of("ATTRIBUTE", a) == ("H5T_STD_U32LE", "SCALAR")
    assert dtype_of("ATTRIBUTE", "Solver Final Time") == ("H5T_IEEE_F64LE", "SCALAR")
    assert dtype_of("ATTRIBUTE", "Stored Layout") == ("H5T_STD_I32LE", "SCALAR")
    assert dtype_of("DATASET", "Number of Elements per Partition")[0] == "H5T_STD_U32LE"
    arrays = {"Node Coordinates": ("H5T_IEEE_F64LE", [60, 3]),
              "Local to Global Node ID": ("H5T_STD_U32LE", [60]),
              "Element Global Node IDs": ("H5T_STD_U32LE", [E, 8]),
              "Local to Global Element IDs": ("H5T_STD_U32LE", [E]),
              "Solution Coefficients": ("H5T_IEEE_F64LE", [E, 8, 5])}
    for name, (t, dims) in arrays.items():
        dt, sp = dtype_of("DATASET", name)
        assert dt == t
        assert f"( {int(np.prod(dims))} )" in sp          # flat, like the reference
        assert dtype_of("DATASET", name + "_offsets")[0] == "H5T_STD_U64LE"
        assert list(_dump(f, name + "_dims").astype(int)) == dims
        assert list(_dump(f, name + "_offsets").astype(int)) == [0]

    xyz = _dump(f, "Node Coordinates").reshape(-1, 3)
    e2n = _dump(f, "Element Global Node IDs").astype(int).reshape(E, 8)
    h = (np.array(hi) - np.array(lo)) / np.array(nelem)
    for e in range(E):
        ix, iy, iz = e % 2, (e // 2) % 3, e // 6
        for n in range(8):
            a, b, c = n & 1, (n >> 1) & 1, n >> 2
            want = np.array(lo) + h * np.array([ix + a, iy + b, iz + c])
            assert np.allclose(xyz[e2n[e, n]], want, rtol=0, atol=1e-14)
    assert list(_dump(f, "Local to Global Element IDs").astype(int)) == list(range(E))


def test_errors(tmp_path):
    f = str(tmp_path / "missing.h5")
    with pytest.raises(h5.DGError) as ei:
        h5.read_header(f)
    assert ei.value.status == 1 and "missing.h5" in str(ei.value)
    g = str(tmp_path / "d.h5")
    h5.write_box(g, (1, 1, 2), (0, 0, 0), (1, 1, 1), _state(2, 1))
    import ctypes as C
    buf = np.zeros(5)
    rc = h5.lib().dg_h5_read_coefficients(g.encode(), buf.ctypes.data_as(C.POINTER(C.c_double)), 5, 0)
    assert rc == 3 and b"too small" in h5.lib().dg_h5_last_error()
    rc = h5.lib().dg_h5_read_coefficients(g.encode(), buf.ctypes.data_as(C.POINTER(C.c_double)), 5, 7)
    assert rc == 3
    with pytest.raises(ValueError):
        h5.write_box(g, (2, 2, 2), (0, 0, 0), (1, 1, 1), _state(3, 1))
    # a text file is not an HDF5 file
    t = tmp_path / "x.h5"
    t.write_text("not hdf5")
    with pytest.raises(h5.DGError):
        h5.read_header(str(t))

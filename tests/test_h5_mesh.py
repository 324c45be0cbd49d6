"""Mesh files on the CPU: the datasets the reference's mesh reader opens
(Mesh::read_mesh, src/mesh/mesh.cpp:17-26,79-200), read from the
reference's own mesh file (examples/quad_2x2/quad_2x2.h5, a 2D quad mesh,
kept as tests/golden/quad_2x2.h5) and written / read back for 3D hex and
2D quadrilateral meshes.  Values are checked against the HDF5 command-line dump (h5dump),
independently of the library."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import dg_amd
import meshgen
from conftest import ROOT

h5 = pytest.importorskip("dg_amd.h5")
if not os.path.exists(h5.H5_LIB_PATH):
    pytest.skip("libdg_amd_h5.so not built (no HDF5 at build time)", allow_module_level=True)

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
QUAD = os.path.join(ROOT, "tests", "golden", "quad_2x2.h5")
CASE = dg_amd.case_toml(order=2, basis="LagrangeHex", timestep=0.01, lower=(0, 0, 0), upper=(1, 1, 1),
                        ic="Uniform", ic_data=(1.0, 0.1, 0.2, 0.3, 2.5), box=False)


def _dump(path, name):
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-d", name, "-y", "-w", "0", path], capture_output=True, text=True,
                         check=True).stdout
    body = out[out.index("DATA {") + 6:]
    body = body[:body.index("}")]
    return np.array([float(x) for x in body.replace(",", " ").split()])


def _types(path):
    head = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    return dict(re.findall(r'DATASET "([^"]+)" \{\s*DATATYPE\s+(\S+)', head))


def test_exports_match_header():
    hdr = open(os.path.join(ROOT, "include", "dg_amd_h5.h")).read()
    syms = re.findall(r"^\s*(?:int|const char\*)\s+(dg_\w+)\s*\(", hdr, re.M)
    assert set(syms) == set(h5.EXPORTS)
    for s in syms:
        assert hasattr(h5.lib(), s)


def test_reads_the_reference_mesh_file():
    info = h5.read_mesh_info(QUAD)
    assert info == {"dim": 2, "qorder": 1, "num_elems": 4, "num_nodes": 9, "num_ifaces": 4,
                    "num_nodes_per_elem": 4, "num_nodes_per_face": 2, "num_bface_groups": 4}
    xyz, e2n, ifc = h5.read_mesh(QUAD)
    # the file's own values (h5dump of examples/quad_2x2/quad_2x2.h5)
    assert e2n.tolist() == [[3, 7, 6, 8], [6, 8, 2, 5], [7, 0, 8, 4], [8, 4, 5, 1]]
    assert ifc.tolist() == [[0, 2, 1, 1, 0, 0], [0, 1, 0, 2, 3, 1], [2, 2, 0, 3, 0, 1], [1, 1, 1, 3, 3, 0]]
    # (h5dump prints the mid-side coordinates rounded; they are 0.5 to 1.3e-12)
    assert np.allclose(xyz, [[0, 0], [1, 0], [1, 1], [0, 1], [0.5, 0], [1, 0.5], [0.5, 1], [0, 0.5], [0.5, 0.5]],
                       rtol=0, atol=1e-11)
    if H5DUMP:
        assert np.array_equal(_dump(QUAD, "NodeCoords"), xyz.ravel())
        assert np.array_equal(_dump(QUAD, "Elem2Nodes"), e2n.ravel())
        assert np.array_equal(_dump(QUAD, "IFaceData"), ifc.ravel())


CASE2 = dg_amd.case_toml(order=2, basis="LagrangeQuad", flux="HLLC", timestep=0.01, lower=(0, 0), upper=(1, 1),
                         ic="Uniform", ic_data=(1.0, 0.1, 0.2, 2.5), box=False)


def test_reference_quad_mesh_file_checks():
    """The file's Dimension must be the case's; the reference's 2x2 quad
    mesh has boundary face groups (walls), which this periodic path
    reports as unsupported."""
    with pytest.raises(dg_amd.DGError) as ei:
        h5.solver_from_mesh_file(CASE, QUAD)
    assert ei.value.status == 1 and "Dimension = 2 does not match Physics.dim = 3" in str(ei.value)
    with pytest.raises(dg_amd.DGError) as ei:
        h5.solver_from_mesh_file(CASE2, QUAD)
    assert ei.value.status == 4 and "boundary face groups" in str(ei.value)


def test_write_read_round_trip_2d(tmp_path):
    import meshgen2d
    from dg_amd import hip
    xyz, e2n, ifc = meshgen2d.box_mesh2d((3, 4), (0, 0), (1, 1), warp=0.1, seed=2, rotate=True, shuffle=True,
                                         random_left=True)
    f = str(tmp_path / "quad.h5")
    h5.write_mesh(f, xyz, e2n, ifc)
    info = h5.read_mesh_info(f)
    assert info == {"dim": 2, "qorder": 1, "num_elems": 12, "num_nodes": 4 * 5, "num_ifaces": 24,
                    "num_nodes_per_elem": 4, "num_nodes_per_face": 2, "num_bface_groups": 0}
    x2, e2, i2 = h5.read_mesh(f)
    assert np.array_equal(x2, xyz) and np.array_equal(e2, e2n) and np.array_equal(i2, ifc)
    if H5DUMP:
        ref, mine = _types(QUAD), _types(f)
        for name, t in mine.items():
            assert ref[name] == t, name
        assert set(mine) == {k for k in ref if not k.startswith("BFG_")}
        assert np.array_equal(_dump(f, "NodeCoords"), xyz.ravel())
        assert np.array_equal(_dump(f, "Elem2Nodes"), e2n.ravel())
    h = C.c_void_p()
    rc = h5.lib().dg_create_from_mesh_file(CASE2.encode(), f.encode(), C.byref(h))
    if rc == 0:
        dg_amd.lib().dg_destroy(h)
    assert rc == (0 if hip.device_count() > 0 else 2), h5.lib().dg_h5_last_error().decode()


def test_write_read_round_trip(tmp_path):
    xyz, e2n, ifc = meshgen.box_mesh((3, 2, 4), (0, 0, 0), (1, 2, 3), warp=0.1, seed=2, rotate=True, shuffle=True,
                                     random_left=True, m=4)
    f = str(tmp_path / "mesh.h5")
    h5.write_mesh(f, xyz, e2n, ifc)
    info = h5.read_mesh_info(f)
    assert info == {"dim": 3, "qorder": 1, "num_elems": 24, "num_nodes": 4 * 3 * 5, "num_ifaces": 72,
                    "num_nodes_per_elem": 8, "num_nodes_per_face": 4, "num_bface_groups": 0}
    x2, e2, i2 = h5.read_mesh(f)
    assert np.array_equal(x2, xyz) and np.array_equal(e2, e2n) and np.array_equal(i2, ifc)
    if H5DUMP:
        # the element types of the reference's own mesh file, dataset by dataset
        ref = _types(QUAD)
        mine = _types(f)
        for name, t in mine.items():
            assert ref[name] == t, name
        assert set(mine) == {k for k in ref if not k.startswith("BFG_")}
        assert np.array_equal(_dump(f, "NodeCoords"), xyz.ravel())
        assert np.array_equal(_dump(f, "IFaceData"), ifc.ravel())


def test_create_from_mesh_file_checks(tmp_path):
    from dg_amd import hip
    want_ok = 0 if hip.device_count() > 0 else 2      # without a GPU the device is the first thing missing
    xyz, e2n, ifc = meshgen.box_mesh((2, 2, 3), (0, 0, 0), (1, 1, 1), warp=0.05, seed=4, rotate=True, m=4)
    f = str(tmp_path / "mesh.h5")
    h5.write_mesh(f, xyz, e2n, ifc)

    def create(text, path):
        h = C.c_void_p()
        rc = h5.lib().dg_create_from_mesh_file(text.encode(), path.encode() if path else None, C.byref(h))
        if rc == 0:
            dg_amd.lib().dg_destroy(h)
        return rc, h5.lib().dg_h5_last_error().decode()

    rc, msg = create(CASE, f)
    assert rc == want_ok, msg
    # Mesh.file of the case, relative to the working directory as in the reference
    cwd = os.getcwd()
    try:
        os.chdir(tmp_path)
        rc, msg = create(CASE.replace("[Mesh]", '[Mesh]\nfile = "mesh.h5"'), None)
        assert rc == want_ok, msg
        rc, msg = create(CASE.replace("[Mesh]", '[Mesh]\nfile = "nothere.h5"'), None)
        assert rc == 1 and "nothere.h5" in msg, msg
    finally:
        os.chdir(cwd)
    # a PeriodicBox case goes to dg_create
    rc, msg = create(dg_amd.case_toml(order=2, nelem=(2, 2, 2), timestep=0.01), None)
    assert rc == want_ok, msg
    # mesh errors surface through dg_h5_last_error
    g = str(tmp_path / "open.h5")
    h5.write_mesh(g, xyz, e2n, ifc[:-2])
    rc, msg = create(CASE, g)
    assert rc == 4 and "boundary faces" in msg, msg
    rc, msg = create(CASE.replace("order = 2", "order = 5"), f)
    assert rc == 4 and "order > 4" in msg, msg


def test_write_mesh_rejects_bad_input(tmp_path):
    xyz, e2n, ifc = meshgen.box_mesh((2, 2, 2), (0, 0, 0), (1, 1, 1), m=3)
    bad = e2n.copy()
    bad[0, 0] = len(xyz)
    with pytest.raises(dg_amd.DGError) as ei:
        h5.write_mesh(str(tmp_path / "m.h5"), xyz, bad, ifc)
    assert ei.value.status == 3
    with pytest.raises(ValueError):
        h5.write_mesh(str(tmp_path / "m.h5"), xyz[:, :2], e2n, ifc)
    with pytest.raises(dg_amd.DGError):
        h5.read_mesh_info(str(tmp_path / "none.h5"))

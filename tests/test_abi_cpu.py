"""CPU checks of the C ABI: the library loads, exports every symbol that
include/dg_amd.h declares, and reports input errors as status codes (the
reference's InputException cases) before touching a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import dg_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "dg_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(dg_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = dg_amd.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    for name in syms:
        assert hasattr(L, name), name
    assert set(syms) == set(dg_amd.EXPORTS)


def _create(text):
    h = C.c_void_p()
    rc = dg_amd.lib().dg_create(text.encode(), C.byref(h))
    msg = dg_amd.lib().dg_last_error(None).decode()
    if rc == 0:
        dg_amd.lib().dg_destroy(h)
    return rc, msg


BASE = dg_amd.case_toml(order=2, nelem=(2, 2, 2), timestep=0.01)


@pytest.mark.parametrize("edit,status,needle", [
    (("convective_flux_fcn = \"Roe\"", "convective_flux_fcn = \"Rusanov\""), 1, "Rusanov is not a valid choice"),
    (("basis = \"LegendreHex\"", "basis = \"LegendreQuad\""), 1, "does not match Physics.dim = 3"),
    (("basis = \"LegendreHex\"", "basis = \"LagrangeSeg\""), 4, "LagrangeSeg"),
    (("basis = \"LegendreHex\"", "basis = \"Spline\""), 1, "not a valid choice"),
    (("type = \"RK4\"", "type = \"SSPRK3\""), 1, "SSPRK3 is not a valid choice"),
    (("dim = 3", "dim = 2"), 1, "does not match Physics.dim = 2"),
    (("dim = 3", "dim = 4"), 1, "Physics.dim must be 2 or 3"),
    (("order = 2", "order = 6"), 4, "no GPU kernel"),
    (("order = 2", "order = 9"), 1, "Legendre polynomial P = 9"),
    (("timestep = 0.01", "timestep = \"x\""), 1, "must be a number"),
    (("name = \"ExtrudedIsentropicVortex\"", "name = \"Gaussian\""), 4, "2D initial condition"),
    (("GasConstant = 1.0", "GasConstant = 287.05"), 1, "GasConstant = 1"),
    (("type = \"PeriodicBox\"", "file = \"meshes/cube_4.h5\""), 4, "HDF5"),
    (("nelem = [2, 2, 2]", "nelem = [2, 0, 2]"), 1, "positive integers"),
    (("data = [1.0, 1.0, 1.0, 1.0, 5.0]", "data = [1.0, 1.0]"), 1, "needs 5 values"),
    (("[Numerics]", "[Numerics\n"), 1, "TOML parse error"),
    (("warp = 0.0", "warp = 0.3"), 1, "Mesh.warp must lie in"),
    (("normals = \"reference\"", "normals = \"average\""), 1, "Mesh.normals"),
])
def test_input_errors(edit, status, needle):
    text = BASE.replace(*edit)
    assert text != BASE
    rc, msg = _create(text)
    assert rc == status, msg
    assert needle in msg, msg


def test_warped_order_limit():
    text = BASE.replace("warp = 0.0", "warp = 0.1").replace("order = 2", "order = 5")
    rc, msg = _create(text)
    assert rc == 4 and "order > 4" in msg, msg


def test_missing_keys():
    rc, msg = _create(BASE.replace("end = 1.0", ""))
    assert rc == 1 and "Stepper.end" in msg
    rc, msg = _create(BASE.replace("timestep = 0.01", ""))
    assert rc == 1 and "Stepper.timestep" in msg


def test_toml_subset_features():
    text = BASE.replace("[Stepper]", "# comment line\n[Stepper]  # trailing\n").replace(
        "data = [1.0, 1.0, 1.0, 1.0, 5.0]", "data = [\n  1.0, 1.0, # rho u\n  1.0, 1.0,\n  5.0,\n]")
    rc, msg = _create(text)
    # parses; only the device may be missing on a CPU-only host
    assert rc in (0, 2), msg


def test_valid_case_without_gpu_fails_loudly():
    from dg_amd import hip
    if hip.device_count() > 0:
        pytest.skip("GPU present")
    rc, msg = _create(BASE)
    assert rc == 2 and "device" in msg.lower()
    with pytest.raises(dg_amd.DGError):
        dg_amd.Solver(BASE)


# ---- dg_create_mesh: mesh arrays in the reference's mesh-file layout ----

MESH_CASE = dg_amd.case_toml(order=2, basis="LagrangeHex", timestep=0.01, lower=(0, 0, 0), upper=(1, 1, 1),
                             ic="Uniform", ic_data=(1.0, 0.1, 0.2, 0.3, 2.5), box=False)


def _mesh():
    import meshgen
    return meshgen.box_mesh((2, 3, 2), (0, 0, 0), (1, 1, 1), warp=0.1, seed=3, rotate=True, shuffle=True,
                            random_left=True, m=4)


def _create_mesh(text, xyz, e2n, ifc):
    import numpy as np
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    e2n = np.ascontiguousarray(e2n, dtype=np.int64)
    ifc = np.ascontiguousarray(ifc, dtype=np.int32)
    h = C.c_void_p()
    rc = dg_amd.lib().dg_create_mesh(text.encode(), len(xyz), xyz.ctypes.data_as(C.POINTER(C.c_double)), len(e2n),
                                     e2n.ctypes.data_as(C.POINTER(C.c_long)), len(ifc),
                                     ifc.ctypes.data_as(C.POINTER(C.c_int)), C.byref(h))
    msg = dg_amd.lib().dg_last_error(None).decode()
    if rc == 0:
        dg_amd.lib().dg_destroy(h)
    return rc, msg


def test_mesh_input_is_checked_before_the_device():
    """A valid mesh gets as far as the device (status 2 without a GPU);
    broken IFaceData, node indices and keys are input errors on any host
    (Mesh<dim>::read_mesh_file's checks, src/mesh/mesh.cpp:79-175)."""
    from dg_amd import hip
    xyz, e2n, ifc = _mesh()
    rc, msg = _create_mesh(MESH_CASE, xyz, e2n, ifc)
    assert rc == (0 if hip.device_count() > 0 else 2), msg
    bad = ifc.copy()
    bad[3, 2] = 9                                   # orientation outside 0..7
    rc, msg = _create_mesh(MESH_CASE, xyz, e2n, bad)
    assert rc == 1 and "orientation 9" in msg, msg
    bad = ifc.copy()
    bad[0, 3] = len(e2n)                            # element out of range
    rc, msg = _create_mesh(MESH_CASE, xyz, e2n, bad)
    assert rc == 1 and "out of range" in msg, msg
    rc, msg = _create_mesh(MESH_CASE, xyz, e2n, ifc[:-1])   # a face without a neighbour
    assert rc == 4 and "boundary faces" in msg, msg
    bad = ifc.copy()
    bad[1] = bad[0]                                 # a face used twice
    rc, msg = _create_mesh(MESH_CASE, xyz, e2n, bad)
    assert rc == 1 and "twice" in msg, msg
    e = e2n.copy()
    e[2, 5] = len(xyz)
    rc, msg = _create_mesh(MESH_CASE, xyz, e, ifc)
    assert rc == 1 and "node index" in msg, msg
    rc, msg = _create_mesh(MESH_CASE.replace("order = 2", "order = 5"), xyz, e2n, ifc)
    assert rc == 4 and "order > 4" in msg, msg
    rc, msg = _create_mesh(MESH_CASE.replace("[Mesh]", "[Mesh]\nwarp = 0.1"), xyz, e2n, ifc)
    assert rc == 1 and "periodic box only" in msg, msg
    rc, msg = _create_mesh(MESH_CASE.replace("timestep = 0.01", "cfl = 0.1"), xyz, e2n, ifc)
    assert rc == 1 and "timestep is required" in msg, msg


BASE2 = dg_amd.case_toml(order=2, basis="LagrangeQuad", flux="HLLC", nelem=(3, 2), lower=(-5, -5), upper=(5, 5),
                         timestep=0.01, ic="IsentropicVortex", ic_data=(1.0, 1.0, 1.0, 1.0, 5.0))


@pytest.mark.parametrize("edit,status,needle", [
    (("nelem = [3, 2]", "nelem = [3, 2, 2]"), 1, "Mesh.nelem must hold 2 integers"),
    (("lower = [-5.0, -5.0]", "lower = [-5.0, -5.0, -5.0]"), 1, "Mesh.lower must hold 2 numbers"),
    (("warp = 0.0", "warp = 0.1"), 4, "dim = 2"),
    (("name = \"IsentropicVortex\"", "name = \"TaylorGreen\""), 4, "3D initial condition"),
    (("data = [1.0, 1.0, 1.0, 1.0, 5.0]", "data = [1.0, 1.0]"), 1, "needs 5 values"),
    (("GasConstant = 1.0", "GasConstant = 287.05"), 1, "GasConstant = 1"),
    (("basis = \"LagrangeQuad\"", "basis = \"LagrangeHex\""), 1, "does not match Physics.dim = 2"),
])
def test_input_errors_2d(edit, status, needle):
    """The dim = 2 case (Physics<2>, quadrilateral bases and ICs,
    src/physics/base/base.cpp:77-88) is checked like the 3D one."""
    text = BASE2.replace(*edit)
    assert text != BASE2
    rc, msg = _create(text)
    assert rc == status, msg
    assert needle in msg, msg


def test_valid_2d_cases_reach_the_device():
    from dg_amd import hip
    want = 0 if hip.device_count() > 0 else 2
    for ic, data in (("Uniform", (1.0, 0.1, 0.2, 2.5)), ("Gaussian", (1.0, 0.5, 0.5, 2.0)),
                     ("IsentropicVortex", (1.0, 1.0, 1.0, 1.0, 5.0))):
        text = BASE2.replace('name = "IsentropicVortex"', f'name = "{ic}"').replace(
            "data = [1.0, 1.0, 1.0, 1.0, 5.0]", "data = [" + ", ".join(repr(float(x)) for x in data) + "]")
        rc, msg = _create(text)
        assert rc == want, msg


MESH_CASE2 = dg_amd.case_toml(order=2, basis="LagrangeQuad", flux="HLLC", timestep=0.01, lower=(0, 0),
                              upper=(1, 1), ic="Uniform", ic_data=(1.0, 0.1, 0.2, 2.5), box=False)


def test_2d_mesh_input_is_checked_before_the_device():
    """Quadrilateral meshes (NodeCoords [n][2], Elem2Nodes [E][4], IFaceData
    with orientations 0..1, src/mesh/mesh.cpp:79-175) get the checks of the
    3D ones."""
    from dg_amd import hip
    import meshgen2d
    xyz, e2n, ifc = meshgen2d.box_mesh2d((3, 2), (0, 0), (1, 1), warp=0.1, seed=4, rotate=True, shuffle=True,
                                         random_left=True)
    rc, msg = _create_mesh(MESH_CASE2, xyz, e2n, ifc)
    assert rc == (0 if hip.device_count() > 0 else 2), msg
    bad = ifc.copy()
    bad[2, 5] = 2                                   # orientation outside 0..1
    rc, msg = _create_mesh(MESH_CASE2, xyz, e2n, bad)
    assert rc == 1 and "not 0..1" in msg, msg
    bad = ifc.copy()
    bad[0, 4] = 4                                   # face id outside 0..3
    rc, msg = _create_mesh(MESH_CASE2, xyz, e2n, bad)
    assert rc == 1 and "out of range" in msg, msg
    rc, msg = _create_mesh(MESH_CASE2, xyz, e2n, ifc[1:])
    assert rc == 4 and "boundary faces" in msg, msg
    bad = ifc.copy()
    bad[1] = bad[0]
    rc, msg = _create_mesh(MESH_CASE2, xyz, e2n, bad)
    assert rc == 1 and "twice" in msg, msg
    e = e2n.copy()
    e[1, 3] = -1
    rc, msg = _create_mesh(MESH_CASE2, xyz, e, ifc)
    assert rc == 1 and "node index" in msg, msg
    with pytest.raises(ValueError, match="2D"):
        dg_amd.Solver(MESH_CASE2, mesh=(xyz, np.zeros((6, 8), dtype=np.int64), ifc))


def test_precision_key():
    """Numerics.precision (a new key): FP64 (default), FP32, BF16 or BF16x3,
    the reduced forms on the affine 3D box only; Stepper.compensated is a
    boolean (3D only)."""
    from dg_amd import hip
    base = dg_amd.case_toml(order=2, nelem=(2, 2, 2), timestep=0.01, precision="FP32")
    rc, msg = _create(base)
    assert rc == (0 if hip.device_count() > 0 else 2), msg
    for prec in ("BF16", "BF16x3"):
        rc, msg = _create(base.replace('precision = "FP32"', f'precision = "{prec}"'))
        assert rc == (0 if hip.device_count() > 0 else 2), msg
    rc, msg = _create(base.replace("timestep = 0.01", "timestep = 0.01\ncompensated = 1"))
    assert rc == 1 and "compensated" in msg, msg
    rc, msg = _create(base.replace("timestep = 0.01", "timestep = 0.01\ncompensated = false"))
    assert rc == (0 if hip.device_count() > 0 else 2), msg
    rc, msg = _create(base.replace('precision = "FP32"', 'precision = "FP16"'))
    assert rc == 1 and "FP16" in msg, msg
    rc, msg = _create(base.replace("warp = 0.0", "warp = 0.1"))
    assert rc == 4 and "precision = FP32 with Mesh.warp" in msg, msg
    # Numerics.subtract_mean: a boolean, on by default for the reduced
    # forms, rejected for FP64
    rc, msg = _create(base.replace('precision = "FP32"', 'precision = "FP32"\nsubtract_mean = false'))
    assert rc == (0 if hip.device_count() > 0 else 2), msg
    rc, msg = _create(base.replace('precision = "FP32"', 'precision = "FP32"\nsubtract_mean = 2'))
    assert rc == 1 and "subtract_mean" in msg, msg
    rc, msg = _create(base.replace('precision = "FP32"', 'subtract_mean = true'))
    assert rc == 4 and "subtract_mean" in msg, msg


def test_precision_fp32_not_on_2d_or_meshes():
    text = BASE2.replace('basis = "LagrangeQuad"', 'basis = "LagrangeQuad"\nprecision = "FP32"')
    rc, msg = _create(text)
    assert rc == 4 and "Physics.dim = 2" in msg, msg
    import meshgen
    xyz, e2n, ifc = meshgen.box_mesh((2, 2, 2), (0, 0, 0), (1, 1, 1), m=3)
    rc, msg = _create_mesh(MESH_CASE.replace('basis = "LagrangeHex"', 'basis = "LagrangeHex"\nprecision = "FP32"'),
                           xyz, e2n, ifc)
    assert rc == 4 and "on a mesh" in msg, msg


def test_batch_create_checks_arguments():
    h = C.c_void_p()
    rc = dg_amd.lib().dg_batch_create(0, None, C.byref(h))
    assert rc == 3 and "cases is NULL" in dg_amd.lib().dg_batch_last_error(None).decode()
    arr = (C.c_void_p * 1)(None)
    rc = dg_amd.lib().dg_batch_create(0, arr, C.byref(h))
    assert rc == 3 and "num_cases < 1" in dg_amd.lib().dg_batch_last_error(None).decode()
    rc = dg_amd.lib().dg_batch_create(1, arr, C.byref(h))
    assert rc == 3 and "case handle is NULL" in dg_amd.lib().dg_batch_last_error(None).decode()
    assert dg_amd.lib().dg_batch_destroy(None) == 0

"""data.h5 solution files with the GPU solver: write (src/exec/main.cpp:
106-112), read back (Solver::read_in_coefficients, src/solver/base.cpp:
233-279) and restart.  Bit-exact: the file stores FP64 coefficients and the
stage kernels are deterministic."""
import os

import numpy as np
import pytest

import dg_amd
from dg_amd import hip

pytestmark = pytest.mark.gpu

h5 = pytest.importorskip("dg_amd.h5")
if not os.path.exists(h5.H5_LIB_PATH):
    pytest.skip("libdg_amd_h5.so not built (no HDF5 at build time)", allow_module_level=True)


def _case(ne=(3, 4, 2), order=2):
    return dg_amd.case_toml(order=order, basis="LegendreHex", flux="Roe", nelem=ne, lower=(-5, -5, -5),
                            upper=(5, 5, 5), stepper="RK4", timestep=0.02, end=1.0,
                            ic="ExtrudedIsentropicVortex", ic_data=(1, 1, 1, 1, 5))


@pytest.fixture(scope="module")
def gpu():
    assert hip.device_count() > 0, "GPU test needs a GPU"


def test_write_then_read_solution(gpu, tmp_path):
    f = str(tmp_path / "data.h5")
    with dg_amd.Solver(_case()) as a, dg_amd.Solver(_case()) as b, dg_amd.Solver(_case()) as c:
        a.init_state_from_fcn()
        a.take_time_step(0.02, 3)
        a.write(f)
        hdr = h5.read_header(f)
        assert (hdr["num_elems"], hdr["num_basis"], hdr["num_state_vars"]) == a.shape
        assert hdr["time"] == a.time and hdr["stored_layout"] == 1
        assert np.array_equal(h5.read_coefficients(f), a.get_state())
        b.read_in_coefficients(f)
        assert np.array_equal(b.get_state(), a.get_state())
        assert b.time == a.time
        assert np.array_equal(b.get_error(), a.get_error())
        # the reference's post keeps its time at 0
        c.read_in_coefficients(f, set_time=False)
        assert c.time == 0.0
        assert np.array_equal(c.get_state(), a.get_state())


def test_restart_is_bit_exact(gpu, tmp_path):
    f = str(tmp_path / "restart.h5")
    with dg_amd.Solver(_case(order=3)) as full, dg_amd.Solver(_case(order=3)) as first, \
            dg_amd.Solver(_case(order=3)) as second:
        full.init_state_from_fcn()
        full.take_time_step(0.02, 6)
        first.init_state_from_fcn()
        first.take_time_step(0.02, 3)
        first.write(f)
        second.read_in_coefficients(f)
        second.take_time_step(0.02, 3)
        assert np.array_equal(second.get_state(), full.get_state())
        assert abs(second.time - full.time) < 1e-15


def test_read_solution_rejects_other_cases(gpu, tmp_path):
    f = str(tmp_path / "data.h5")
    with dg_amd.Solver(_case()) as a, dg_amd.Solver(_case(ne=(3, 4, 3))) as b, \
            dg_amd.Solver(_case(order=3)) as c:
        a.init_state_from_fcn()
        a.write(f)
        for s in (b, c):
            with pytest.raises(dg_amd.DGError) as ei:
                s.read_in_coefficients(f)
            assert ei.value.status == 1 and "the case has" in str(ei.value)
        with pytest.raises(dg_amd.DGError):
            b.read_in_coefficients(str(tmp_path / "missing.h5"))


def test_file_written_in_column_major_reads_back(gpu, tmp_path):
    """A file with "Stored Layout" = 0 (the reference's LayoutLeft GPU
    builds, src/solver/base.cpp:478-489) loads into the same state."""
    f = str(tmp_path / "left.h5")
    with dg_amd.Solver(_case()) as a, dg_amd.Solver(_case()) as b:
        a.init_state_from_fcn()
        U = a.get_state()
        i = a.info
        h5.write_box(f, tuple(i.nelem), tuple(i.lower), tuple(i.upper),
                     np.ascontiguousarray(U.transpose(2, 1, 0)), time=0.0, layout=dg_amd.DG_LAYOUT_LEFT)
        b.read_in_coefficients(f)
        assert np.array_equal(b.get_state(), U)


# ---- mesh files (src/mesh/mesh.cpp:79-200) and data.h5 of mesh cases ----

import json  # noqa: E402
import re  # noqa: E402
import shutil  # noqa: E402
import subprocess  # noqa: E402

import meshgen  # noqa: E402

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
BINDIR = os.path.join(os.path.dirname(dg_amd.PKG_ROOT), "dg-mpi_amd", "bin")


def _mesh_case(order=3, normals="reference"):
    return dg_amd.case_toml(order=order, basis="LagrangeHex", flux="HLLC", lower=(0, 0, 0), upper=(1, 1, 1),
                            timestep=0.002, end=1.0, ic="IsentropicVortex", ic_data=(1, 1, 1, 1, 5),
                            normals=normals, box=False)


def _dump(path, name):
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-d", name, "-y", "-w", "0", path], capture_output=True,
                         text=True, check=True).stdout
    body = out[out.index("DATA {") + 6:]
    return np.array([float(x) for x in body[:body.index("}")].replace(",", " ").split()])


def test_solver_from_mesh_file_matches_the_arrays(gpu, tmp_path):
    mesh = meshgen.box_mesh((3, 3, 2), (0, 0, 0), (1, 1, 1), warp=0.1, seed=9, rotate=True, shuffle=True,
                            random_left=True, m=5)
    f = str(tmp_path / "mesh.h5")
    h5.write_mesh(f, *mesh)
    with h5.solver_from_mesh_file(_mesh_case(), f) as a, dg_amd.Solver(_mesh_case(), mesh=mesh) as b:
        xyz, e2n = a.mesh()
        assert np.array_equal(xyz, mesh[0]) and np.array_equal(e2n, mesh[1])
        for s in (a, b):
            s.init_state_from_fcn()
            s.take_time_step(0.0, 3)
        assert np.array_equal(a.get_state(), b.get_state())


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_solution_file_carries_the_solver_mesh(gpu, tmp_path):
    """data.h5 stores the solver's own nodes: a given mesh, and the warped
    box's displaced grid (Writer::Writer, src/io/writer.cpp:68-98)."""
    mesh = meshgen.box_mesh((2, 3, 2), (0, 0, 0), (1, 1, 1), warp=0.08, seed=1, rotate=True, shuffle=True, m=5)
    warped = dg_amd.case_toml(order=2, basis="LagrangeHex", flux="Roe", nelem=(3, 3, 4), lower=(0, 0, 0),
                              upper=(1, 1, 1), timestep=0.002, ic="IsentropicVortex", ic_data=(1, 1, 1, 1, 5),
                              warp=0.1)
    for k, s in enumerate([dg_amd.Solver(_mesh_case(), mesh=mesh), dg_amd.Solver(warped)]):
        with s:
            s.init_state_from_fcn()
            f = str(tmp_path / f"data{k}.h5")
            s.write(f)
            xyz, e2n = s.mesh()
            assert np.array_equal(_dump(f, "Node Coordinates"), xyz.ravel())
            assert np.array_equal(_dump(f, "Element Global Node IDs"), e2n.ravel())
            if k == 0:
                assert np.array_equal(xyz, mesh[0]) and np.array_equal(e2n, mesh[1])
            else:
                # the box grid with the periodic displacement: interior nodes moved
                lo = np.array([[i / 3, j / 3, kk / 4] for kk in range(5) for j in range(4) for i in range(4)])
                assert xyz.shape == lo.shape and np.max(np.abs(xyz - lo)) > 1e-3
            assert np.array_equal(h5.read_coefficients(f), s.get_state())


def test_dg_run_on_the_reference_example_with_a_mesh_file(gpu, tmp_path):
    """The reference's examples/euler/3D/isentropic_vortex/input.toml,
    unchanged (Mesh.file = "meshes/cube_4.h5"), on a periodic 4^3 cube mesh
    file written in the reference's mesh layout (elements in their own
    frames); dg_post's errors match the example's regression data."""
    gd = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "vortex3d_regression.json")))
    (tmp_path / "meshes").mkdir()
    h5.write_mesh(str(tmp_path / "meshes" / "cube_4.h5"),
                  *meshgen.box_mesh((4, 4, 4), (0, 0, 0), (1, 1, 1), seed=21, rotate=True, shuffle=True,
                                    random_left=True, m=5))
    (tmp_path / "input.toml").write_text("""[Mesh]
file = "meshes/cube_4.h5"

[Numerics]
order = 3
basis = "LagrangeHex"

[Stepper]
start = 0.0
end = 1.0
timestep = 1e-3
type = "RK4"

[Physics]
name = "Euler"
dim = 3
gamma = 1.4
GasConstant = 1.0
convective_flux_fcn = "HLLC"

[InitialCondition]
name = "IsentropicVortex"
data = [1.0, 1.0, 1.0, 1.0, 5.0]
""")
    r = subprocess.run([os.path.join(BINDIR, "dg_run")], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Number of elements = 64" in r.stdout and "Number of iterations = 1000" in r.stdout
    r = subprocess.run([os.path.join(BINDIR, "dg_post")], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    errs = [float(v) for v in re.findall(r"norm_total_error\(\d\)=([-+0-9.eE]+)", r.stdout)]
    ref = [float(v) for v in gd["l2_errors"]]
    assert len(errs) == 5
    for v in (0, 1, 3, 4):
        assert abs(errs[v] - ref[v]) < 1e-10 * ref[v]

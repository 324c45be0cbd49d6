"""GPU parity of 2D quadrilateral meshes (Physics.dim = 2 with NodeCoords
[n][2], Elem2Nodes [E][4], IFaceData [F][6], src/mesh/mesh.cpp:79-175):
renumbered elements in any of the four frames, random face orientations,
bilinear (non-affine) elements, against the 2D mesh oracle
(oracle/dg2d.py Mesh2D).

Bar: 1e-12 relative (max-norm) for the residual; 1e-11 for results through
the per-element inverse mass matrices and for five steps; free stream to
1e-13 absolute."""
import os

import numpy as np
import pytest

import dg_amd
import dg2d
from dg_amd import hip
from meshgen2d import box_mesh2d

pytestmark = pytest.mark.gpu

VORTEX = ("IsentropicVortex", (1.0, 1.0, 1.0, 1.0, 5.0))
LO, HI = (-5, -5), (5, 5)
CASES = [
    # order, basis, flux, nelem, warp, seed
    (1, "LagrangeQuad", "LaxFriedrichs", (3, 4), 0.0, 1),
    (2, "LegendreQuad", "Roe", (4, 3), 0.15, 2),
    (3, "LagrangeQuad", "HLLC", (3, 3), 0.2, 3),
    (4, "LegendreQuad", "HLLC", (2, 3), 0.1, 4),
    (5, "LagrangeQuad", "Roe", (3, 2), 0.1, 5),
]


def case_id(c):
    return f"p{c[0]}-{c[1][:4]}-{c[2]}-{c[3][0]}x{c[3][1]}-w{c[4]}"


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def toml(order, basis, flux, ic=VORTEX, timestep=0.002, stepper="RK4", lo=LO, hi=HI):
    return dg_amd.case_toml(order=order, basis=basis, flux=flux, lower=lo, upper=hi, stepper=stepper,
                            timestep=timestep, ic=ic[0], ic_data=ic[1], box=False)


def make_pair(case, stepper="RK4"):
    order, basis, flux, ne, warp, seed = case
    mesh = box_mesh2d(ne, LO, HI, warp=warp, seed=seed, rotate=True, shuffle=True, random_left=True)
    g = dg_amd.Solver(toml(order, basis, flux, stepper=stepper), mesh=mesh)
    o = dg2d.Mesh2D(order, basis, flux, mesh, LO, HI, ic=VORTEX[0], ic_data=VORTEX[1])
    return g, o, mesh


@pytest.fixture(scope="module")
def gpu():
    assert hip.device_count() > 0, "GPU test needs a GPU"


@pytest.mark.parametrize("case", CASES, ids=case_id)
def test_2d_mesh_projection_residual_inv_mass(gpu, case):
    print(f"case {case_id(case)}", flush=True)
    g, o, _ = make_pair(case)
    assert g.shape == (o.E, o.nb, 4) and g.info.dim == 2
    g.init_state_from_fcn()
    o.init_state()
    assert relerr(g.get_state(), o.U) < 1e-11
    U = o.U
    Ud = hip.Buffer.from_array(U)
    Rd = hip.Buffer(U.size)
    g.get_residual(Ud.addr, Rd.addr)
    hip.synchronize()
    R = o.residual(U)
    assert relerr(Rd.download(U.shape), R) < 1e-12
    g.mult_inv_mass_matrix(0.37, Rd.addr, Ud.addr)
    hip.synchronize()
    assert relerr(Ud.download(U.shape), o.inv_mass(0.37, R)) < 1e-11
    for t in (0.0, 0.3):
        assert np.allclose(g.get_error(t)[:4], o.l2_error(t), rtol=1e-11, atol=1e-15)
    assert np.allclose(g.integrals()[:4], o.integrals(), rtol=1e-11, atol=1e-12)
    g.close()


@pytest.mark.parametrize("stepper", ["RK4", "FE"])
@pytest.mark.parametrize("case", [CASES[1], CASES[2]], ids=case_id)
def test_2d_mesh_time_steps(gpu, case, stepper):
    print(f"case {case_id(case)} {stepper}", flush=True)
    g, o, _ = make_pair(case, stepper=stepper)
    o.init_state()
    g.set_state(o.U)
    g.take_time_step(0.0, 5)
    o.step(stepper, 0.002, 5)
    assert relerr(g.get_state(), o.U) < 1e-11
    g.close()


def test_2d_box_as_a_mesh_is_the_box(gpu):
    """Identity numbering and frames: the mesh path gives the box path's
    state after steps (same arithmetic up to the metric terms' rounding)."""
    ne = (4, 3)
    with dg_amd.Solver(toml(3, "LagrangeQuad", "HLLC"), mesh=box_mesh2d(ne, LO, HI)) as gm, \
            dg_amd.Solver(dg_amd.case_toml(order=3, basis="LagrangeQuad", flux="HLLC", nelem=ne, lower=LO, upper=HI,
                                           timestep=0.002, ic=VORTEX[0], ic_data=VORTEX[1])) as gb:
        for g in (gm, gb):
            g.init_state_from_fcn()
            g.take_time_step(0.0, 5)
        assert relerr(gm.get_state(), gb.get_state()) < 1e-12


@pytest.mark.parametrize("flux", ["LaxFriedrichs", "HLLC", "Roe"])
def test_2d_warped_mesh_free_stream_and_conservation(gpu, flux):
    mesh = box_mesh2d((4, 3), (0, 0), (4, 3), warp=0.15, seed=5, rotate=True, random_left=True)
    with dg_amd.Solver(toml(3, "LegendreQuad", flux, ic=("Uniform", (1.1, 0.4, -0.3, 2.9)), timestep=0.01,
                            lo=(0, 0), hi=(4, 3)), mesh=mesh) as g:
        g.init_state_from_fcn()
        U0 = g.get_state()
        g.take_time_step(0.0, 3)
        assert np.max(np.abs(g.get_state() - U0)) < 1e-13
    with dg_amd.Solver(toml(2, "LagrangeQuad", flux, timestep=0.01), mesh=box_mesh2d((4, 4), LO, HI, warp=0.2)) as g:
        g.init_state_from_fcn()
        I0 = g.integrals()
        g.take_time_step(0.0, 10)
        assert np.allclose(g.integrals()[:4], I0[:4], rtol=1e-13, atol=1e-12)


def test_2d_mesh_with_self_paired_faces(gpu):
    mesh = box_mesh2d((1, 3), LO, HI, warp=0.1, seed=7, rotate=True, random_left=True)
    assert any(r[0] == r[3] for r in mesh[2])
    with dg_amd.Solver(toml(2, "LagrangeQuad", "Roe"), mesh=mesh) as g:
        o = dg2d.Mesh2D(2, "LagrangeQuad", "Roe", mesh, LO, HI, ic=VORTEX[0], ic_data=VORTEX[1])
        o.init_state()
        U = o.U
        Ud = hip.Buffer.from_array(U)
        Rd = hip.Buffer(U.size)
        g.get_residual(Ud.addr, Rd.addr)
        hip.synchronize()
        assert relerr(Rd.download(U.shape), o.residual(U)) < 1e-12
        g.set_state(U)
        g.take_time_step(0.0, 3)
        o.step("RK4", 0.002, 3)
        assert relerr(g.get_state(), o.U) < 1e-11


def test_2d_mesh_file_and_solution_file(gpu, tmp_path):
    """A 2D mesh file (the reference reader's datasets) gives the solver of
    the same arrays; data.h5 carries the mesh's own nodes."""
    h5 = pytest.importorskip("dg_amd.h5")
    if not os.path.exists(h5.H5_LIB_PATH):
        pytest.skip("libdg_amd_h5.so not built")
    case = CASES[2]
    g, o, mesh = make_pair(case)
    f = str(tmp_path / "quad.h5")
    h5.write_mesh(f, *mesh)
    gf = h5.solver_from_mesh_file(toml(*case[:3]), f)
    for s in (g, gf):
        s.init_state_from_fcn()
        s.take_time_step(0.0, 4)
    assert np.array_equal(g.get_state(), gf.get_state())
    xyz, e2n = gf.mesh()
    assert np.array_equal(xyz, mesh[0]) and np.array_equal(e2n, mesh[1])
    out = str(tmp_path / "data.h5")
    gf.write(out)
    hdr = h5.read_header(out)
    assert (hdr["dim"], hdr["num_nodes_per_elem"], hdr["num_elems"], hdr["num_nodes"]) == (2, 4, o.E, len(mesh[0]))
    g2 = h5.solver_from_mesh_file(toml(*case[:3]), f)
    g2.read_in_coefficients(out)
    assert np.array_equal(g2.get_state(), gf.get_state())
    for s in (g, gf, g2):
        s.close()

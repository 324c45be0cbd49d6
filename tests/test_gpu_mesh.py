"""GPU parity on hex meshes given in the reference's mesh-file layout
(Mesh<dim>::read_mesh_file, src/mesh/mesh.cpp:79-175): elements renumbered
and each in its own local frame (any of the 24 rotations of the cube), left
orientations drawn at random, non-affine (warped) or affine nodes.  The
face pairing (construct_face_states, src/solver/base.cpp:287-327), the
normals (src/numerics/basis/basis.cpp:320-350) and the per-element mass
matrices follow the oracle run on the same arrays.

Bar: 1e-12 relative (max-norm) for the residual; 1e-11 for results that go
through M^-1 (Gauss-Jordan on the GPU, LU in the oracle) and for five
steps."""
import numpy as np
import pytest

import dg_amd
import meshgen
import oracle as O
from dg_amd import hip

pytestmark = pytest.mark.gpu

TAU = 6.283185307179586
VORTEX = ("ExtrudedIsentropicVortex", (1, 1, 1, 1, 5))
CASES = [
    # order, basis, flux, nelem, lo, hi, (ic, data), warp, normals, mesh options
    (1, "LagrangeHex", "LaxFriedrichs", (3, 4, 3), (-1, -1.5, 0), (2, 1.5, 1), VORTEX, 0.12, "reference",
     dict(rotate=True, shuffle=True, random_left=True, seed=1)),
    (2, "LegendreHex", "Roe", (3, 3, 4), (-5, -5, -5), (5, 5, 5), VORTEX, 0.1, "exact",
     dict(rotate=True, seed=2)),
    (3, "LagrangeHex", "HLLC", (2, 3, 2), (0, 0, 0), (1, 1, 1), ("IsentropicVortex", (1, 1, 1, 1, 5)), 0.0,
     "reference", dict(rotate=True, shuffle=True, random_left=True, seed=3)),
    (3, "LegendreHex", "Roe", (3, 3, 3), (0, 0, 0), (TAU, TAU, TAU), ("TaylorGreen", (0.1,)), 0.1, "exact",
     dict(rotate=True, shuffle=True, random_left=True, seed=4)),
    (4, "LegendreHex", "HLLC", (3, 3, 3), (-5, -5, -5), (5, 5, 5),
     ("GaussianBump", (0.2, 1.5, 0.6, -0.48, 0.64, 1.0, 1.0, 0.0, 2.0)), 0.1, "reference",
     dict(rotate=True, shuffle=True, random_left=True, seed=5)),
]


def case_id(c):
    return f"p{c[0]}-{c[1][:4]}-{c[2]}-{c[6][0][:8]}-w{c[7]}-{c[8]}"


@pytest.fixture(scope="module")
def gpu():
    assert hip.device_count() > 0, "GPU test needs a GPU"


def make_pair(case, stepper="RK4", timestep=0.01):
    order, basis, flux, ne, lo, hi, (ic, data), warp, normals, opts = case
    mesh = meshgen.box_mesh(ne, lo, hi, warp=warp, m=order + 2, **opts)
    toml = dg_amd.case_toml(order=order, basis=basis, flux=flux, lower=lo, upper=hi, stepper=stepper,
                            timestep=timestep, end=1.0, ic=ic, ic_data=data, normals=normals, box=False)
    g = dg_amd.Solver(toml, mesh=mesh)
    o = O.Solver(O.make_params(order=order, basis=basis, flux=flux, nelem=ne, lo=lo, hi=hi, gamma=1.4, R=1.0,
                               ic=ic, ic_data=data, exact_normals=(normals == "exact")), mesh=mesh)
    return g, o


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("case", CASES, ids=case_id)
def test_mesh_init_state_and_residual(gpu, case):
    print(f"case {case_id(case)}", flush=True)
    g, o = make_pair(case)
    assert g.shape == o.shape()
    g.init_state_from_fcn()
    o.init_state()
    U = o.get_state()
    assert relerr(g.get_state(), U) < 1e-11
    Ud = hip.Buffer.from_array(U)
    Rd = hip.Buffer(U.size)
    g.get_residual(Ud.addr, Rd.addr)
    hip.synchronize()
    R = o.residual(U)
    assert relerr(Rd.download(U.shape), R) < 1e-12
    g.mult_inv_mass_matrix(0.37, Rd.addr, Ud.addr)
    hip.synchronize()
    assert relerr(Ud.download(U.shape), o.inv_mass(0.37, R)) < 1e-11
    for t in (0.0, 0.2):
        assert np.allclose(g.get_error(t), o.l2_error(t), rtol=1e-11, atol=1e-15)
    g.close()


@pytest.mark.parametrize("stepper", ["RK4", "FE"])
@pytest.mark.parametrize("case", [CASES[0], CASES[3]], ids=case_id)
def test_mesh_time_steps(gpu, case, stepper):
    dt = 0.002
    print(f"case {case_id(case)} {stepper}", flush=True)
    g, o = make_pair(case, stepper=stepper, timestep=dt)
    o.init_state()
    g.set_state(o.get_state())
    g.take_time_step(0.0, 5)
    o.step(stepper, dt, 5)
    assert relerr(g.get_state(), o.get_state()) < 1e-11
    g.close()


def test_identity_mesh_matches_the_periodic_box(gpu):
    """The warped box passed as a mesh (own element order, unrotated
    frames) takes the same path as the PeriodicBox case: same residual."""
    ne, lo, hi = (3, 4, 3), (-1, -1.5, 0), (2, 1.5, 1)
    mesh = meshgen.box_mesh(ne, lo, hi, warp=0.12, m=5)
    kw = dict(order=3, basis="LagrangeHex", flux="Roe", lower=lo, upper=hi, timestep=0.01, ic=VORTEX[0],
              ic_data=VORTEX[1])
    with dg_amd.Solver(dg_amd.case_toml(box=False, **kw), mesh=mesh) as a, \
            dg_amd.Solver(dg_amd.case_toml(nelem=ne, warp=0.12, **kw)) as b:
        a.init_state_from_fcn()
        b.init_state_from_fcn()
        U = b.get_state()
        assert relerr(a.get_state(), U) < 1e-13
        a.set_state(U)
        a.take_time_step(0.0, 3)
        b.take_time_step(0.0, 3)
        assert relerr(a.get_state(), b.get_state()) < 1e-13


@pytest.mark.parametrize("flux", ["LaxFriedrichs", "HLLC", "Roe"])
def test_rotated_frames_do_not_change_the_solution(gpu, flux):
    """With normals at the flux points the scheme does not depend on the
    elements' local frames or numbering: integrals and the L2 error after
    steps on a rotated, renumbered mesh match the plain one to rounding."""
    ne, lo, hi = (3, 3, 3), (-5, -5, -5), (5, 5, 5)
    kw = dict(order=2, basis="LagrangeHex", flux=flux, lower=lo, upper=hi, timestep=0.02, ic=VORTEX[0],
              ic_data=VORTEX[1], normals="exact", box=False)
    plain = meshgen.box_mesh(ne, lo, hi, warp=0.1, m=4)
    turned = meshgen.box_mesh(ne, lo, hi, warp=0.1, m=4, rotate=True, shuffle=True, random_left=True, seed=11)
    with dg_amd.Solver(dg_amd.case_toml(**kw), mesh=plain) as a, \
            dg_amd.Solver(dg_amd.case_toml(**kw), mesh=turned) as b:
        for s in (a, b):
            s.init_state_from_fcn()
            s.take_time_step(0.0, 4)
        assert np.allclose(a.integrals()[:5], b.integrals()[:5], rtol=1e-12, atol=1e-12)
        assert np.allclose(a.get_error(), b.get_error(), rtol=1e-10, atol=1e-14)


def test_mesh_free_stream_and_conservation(gpu):
    """Uniform flow on a rotated warped mesh stays steady with exact normals;
    the conserved integrals stay put with either normal placement."""
    ne, lo, hi = (3, 4, 3), (-1, -1.5, 0), (2, 1.5, 1)
    mesh = meshgen.box_mesh(ne, lo, hi, warp=0.12, m=5, rotate=True, shuffle=True, random_left=True, seed=8)
    for normals in ("exact", "reference"):
        toml = dg_amd.case_toml(order=3, basis="LagrangeHex", flux="HLLC", lower=lo, upper=hi, timestep=0.01,
                                ic="Uniform", ic_data=(1.2, 0.3, -0.2, 0.5, 3.1), normals=normals, box=False)
        with dg_amd.Solver(toml, mesh=mesh) as g:
            g.init_state_from_fcn()
            U0 = g.get_state()
            I0 = g.integrals()
            assert I0[0] == pytest.approx(1.2 * 9.0, rel=1e-13)
            g.take_time_step(0.0, 3)
            if normals == "exact":
                assert np.max(np.abs(g.get_state() - U0)) < 1e-12
            assert np.allclose(g.integrals()[:5], I0[:5], rtol=1e-12, atol=1e-13)

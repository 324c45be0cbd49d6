"""Edge cases of the connectivity on the GPU: elements that are their own
neighbour across a periodic direction (one element along it), in the box,
in a mesh given as arrays (IFaceData rows pairing two faces of the same
element) and in 2D; against the oracles at 1e-12 (residual) and 1e-11
(three steps)."""
import numpy as np
import pytest

import dg_amd
import dg2d
import meshgen
import oracle as O
from dg_amd import hip

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def residual_of(g, U):
    Ud = hip.Buffer.from_array(U)
    Rd = hip.Buffer(U.size)
    g.get_residual(Ud.addr, Rd.addr)
    hip.synchronize()
    return Rd.download(U.shape)


def test_box_with_one_element_along_x():
    ic = ("ExtrudedIsentropicVortex", (1.0, 1.0, 1.0, 1.0, 5.0))
    with dg_amd.Solver(dg_amd.case_toml(order=2, basis="LagrangeHex", flux="LaxFriedrichs", nelem=(1, 2, 3),
                                        timestep=0.01, ic=ic[0], ic_data=ic[1])) as g:
        o = O.Solver(O.make_params(order=2, basis="LagrangeHex", flux="LaxFriedrichs", nelem=(1, 2, 3),
                                   lo=(-5, -5, -5), hi=(5, 5, 5), ic=ic[0], ic_data=ic[1]))
        o.init_state()
        U = o.get_state()
        assert relerr(residual_of(g, U), o.residual(U)) < 1e-12
        g.set_state(U)
        g.take_time_step(0.0, 3)
        o.step("RK4", 0.01, 3)
        assert relerr(g.get_state(), o.get_state()) < 1e-11


def test_mesh_with_self_paired_faces():
    mesh = meshgen.box_mesh((1, 1, 2), (0, 0, 0), (1, 1, 1), seed=3, rotate=True, random_left=True, m=4)
    assert any(r[0] == r[3] for r in mesh[2])
    toml = dg_amd.case_toml(order=2, basis="LagrangeHex", flux="HLLC", lower=(0, 0, 0), upper=(1, 1, 1),
                            timestep=0.001, ic="IsentropicVortex", ic_data=(), box=False)
    with dg_amd.Solver(toml, mesh=mesh) as g:
        o = O.Solver(O.make_params(order=2, basis="LagrangeHex", flux="HLLC", nelem=(1, 1, 2), lo=(0, 0, 0),
                                   hi=(1, 1, 1), ic="IsentropicVortex", ic_data=()), mesh=mesh)
        o.init_state()
        U = o.get_state()
        assert relerr(residual_of(g, U), o.residual(U)) < 1e-12
        g.set_state(U)
        g.take_time_step(0.0, 3)
        o.step("RK4", 0.001, 3)
        assert relerr(g.get_state(), o.get_state()) < 1e-11


def test_2d_box_with_one_element_along_x():
    with dg_amd.Solver(dg_amd.case_toml(order=2, basis="LagrangeQuad", flux="HLLC", nelem=(1, 3),
                                        lower=(-5, -5), upper=(5, 5), timestep=0.01, ic="IsentropicVortex",
                                        ic_data=(1.0, 1.0, 1.0, 1.0, 5.0))) as g:
        o = dg2d.Box2D(2, "LagrangeQuad", "HLLC", (1, 3), (-5, -5), (5, 5))
        o.init_state()
        assert relerr(residual_of(g, o.U), o.residual(o.U)) < 1e-12
        g.set_state(o.U)
        g.take_time_step(0.0, 3)
        o.step("RK4", 0.01, 3)
        assert relerr(g.get_state(), o.U) < 1e-11


@pytest.mark.parametrize("precision", [None, "FP32"])
def test_box_past_32_bit_offsets(precision):
    """192^3 elements at p = 3 (7.1 M elements, 2.3e9 values per register,
    18 GB each): element offsets e (p+1)^3 5 pass 2^31, so every address of
    the stage, the initial projection and the reductions must be formed in
    64 bits.  The z-extruded vortex on [-15, 15]^3 (the headline's element
    size), two RK4 steps at CFL 0.1: mass, momentum and energy conserved
    (1e-12 of the largest integral, FP32 form 1e-6) and the L2 error against
    the exact vortex within three times its t = 0 projection error -- an
    element read or written at a wrapped offset would put O(1) errors into
    the free stream beyond it."""
    ne = 192
    toml = dg_amd.case_toml(order=3, basis="LegendreHex", flux="Roe", nelem=(ne, ne, ne), lower=(-15.,) * 3,
                            upper=(15.,) * 3, stepper="RK4", cfl=0.1, end=1.0, ic="ExtrudedIsentropicVortex",
                            ic_data=(1.0, 1.0, 1.0, 1.0, 5.0), precision=precision)
    print(f"case 192^3 p3 Roe vortex {precision or 'FP64'}", flush=True)
    with dg_amd.Solver(toml) as g:
        assert g.shape[0] * g.shape[1] * g.shape[2] > 2 ** 31
        g.init_state_from_fcn()
        e0 = np.array(g.get_error())
        I0 = np.array(g.integrals())
        g.take_time_step(g.timestep, 2)
        e = np.array(g.get_error(g.time))
        I1 = np.array(g.integrals())
    print(f"L2 error t=0 {e0}, after 2 steps {e}; integrals {I0[:5]} -> {I1[:5]}", flush=True)
    assert np.all(np.isfinite(e)) and np.all(np.isfinite(I1))
    tol = 1e-12 if precision is None else 1e-6
    assert np.all(np.abs(I1[:5] - I0[:5]) < tol * np.max(np.abs(I0[:5])))
    v = [0, 1, 2, 4]
    assert np.all(e[v] < 3.0 * e0[v] + 1e-12)

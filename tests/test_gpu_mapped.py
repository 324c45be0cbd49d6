"""GPU parity on the warped (non-affine, trilinear) periodic box: per-point
metric terms, general-normal Riemann solvers, per-element dense inverse
mass matrices (src/solver/helpers.cpp:4-100,380-500,
src/solver/flux_functors_impl.h:95-102, src/solver/tools.cpp:124-139),
against the CPU oracle with the same node displacement.

Bar: 1e-12 relative (max-norm) for the residual; 1e-11 for results that go
through M^-1 (the GPU inverts each mass matrix by Gauss-Jordan, the oracle
by LU; two inversions of a matrix agree to about cond(M) eps, cond(M) up to
~1e3 for these elements) and for five steps."""
import numpy as np
import pytest

import dg_amd
import oracle as O
from dg_amd import hip

pytestmark = pytest.mark.gpu

TAU = 6.283185307179586
CASES = [
    # order, basis, flux, nelem, lo, hi, ic, data, warp, normals
    (1, "LagrangeHex", "LaxFriedrichs", (3, 4, 3), (-1, -1.5, 0), (2, 1.5, 1), "ExtrudedIsentropicVortex",
     (1, 1, 1, 1, 5), 0.12, "reference"),
    (2, "LegendreHex", "Roe", (4, 3, 3), (-5, -5, -5), (5, 5, 5), "ExtrudedIsentropicVortex", (1, 1, 1, 1, 5),
     0.1, "exact"),
    (3, "LagrangeHex", "HLLC", (3, 3, 4), (0, 0, 0), (1, 1, 1), "IsentropicVortex", (1, 1, 1, 1, 5), 0.08,
     "reference"),
    (3, "LegendreHex", "Roe", (3, 3, 3), (0, 0, 0), (TAU, TAU, TAU), "TaylorGreen", (0.1,), 0.1, "exact"),
    (4, "LegendreHex", "HLLC", (3, 3, 3), (-5, -5, -5), (5, 5, 5), "GaussianBump",
     (0.2, 1.5, 0.6, -0.48, 0.64, 1.0, 1.0, 0.0, 2.0), 0.1, "reference"),
]


def case_id(c):
    return f"p{c[0]}-{c[1][:4]}-{c[2]}-{c[6][:8]}-{c[9]}"


@pytest.fixture(scope="module")
def gpu():
    assert hip.device_count() > 0, "GPU test needs a GPU"


def make_pair(case, stepper="RK4", timestep=0.01):
    order, basis, flux, ne, lo, hi, ic, data, warp, normals = case
    g = dg_amd.Solver(dg_amd.case_toml(order=order, basis=basis, flux=flux, nelem=ne, lower=lo, upper=hi,
                                       stepper=stepper, timestep=timestep, end=1.0, ic=ic, ic_data=data,
                                       warp=warp, normals=normals))
    o = O.Solver(O.make_params(order=order, basis=basis, flux=flux, nelem=ne, lo=lo, hi=hi, gamma=1.4, R=1.0,
                               ic=ic, ic_data=data, warp=warp, exact_normals=(normals == "exact")))
    return g, o


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("case", CASES, ids=case_id)
def test_mapped_init_state_and_residual(gpu, case):
    g, o = make_pair(case)
    print(f"case {case_id(case)}", flush=True)
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
    # dense inverse mass matrices
    g.mult_inv_mass_matrix(0.37, Rd.addr, Ud.addr)
    hip.synchronize()
    assert relerr(Ud.download(U.shape), o.inv_mass(0.37, R)) < 1e-11
    # L2 error and total volume
    for t in (0.0, 0.2):
        assert np.allclose(g.get_error(t), o.l2_error(t), rtol=1e-11, atol=1e-15)
    g.close()


@pytest.mark.parametrize("stepper", ["RK4", "FE"])
@pytest.mark.parametrize("case", [CASES[0], CASES[1], CASES[4]], ids=case_id)
def test_mapped_time_steps(gpu, case, stepper):
    dt = 0.002
    g, o = make_pair(case, stepper=stepper, timestep=dt)
    print(f"case {case_id(case)} {stepper}", flush=True)
    o.init_state()
    g.set_state(o.get_state())
    g.take_time_step(0.0, 5)
    o.step(stepper, dt, 5)
    assert relerr(g.get_state(), o.get_state()) < 1e-11
    g.close()


@pytest.mark.parametrize("flux", ["LaxFriedrichs", "HLLC", "Roe"])
def test_mapped_free_stream_with_exact_normals(gpu, flux):
    """Uniform flow stays steady on the warped mesh with normals at the flux
    points; the conserved integrals (w |J| on the (2p+2)^3 rule) stay put
    to rounding in both normal placements."""
    for normals in ("exact", "reference"):
        g = dg_amd.Solver(dg_amd.case_toml(order=3, basis="LagrangeHex", flux=flux, nelem=(3, 4, 3),
                                           lower=(-1, -1.5, 0), upper=(2, 1.5, 1), stepper="RK4", timestep=0.01,
                                           ic="Uniform", ic_data=(1.2, 0.3, -0.2, 0.5, 3.1), warp=0.12,
                                           normals=normals))
        g.init_state_from_fcn()
        U0 = g.get_state()
        I0 = g.integrals()
        assert I0[0] == pytest.approx(1.2 * 9.0, rel=1e-13)     # box volume 3 x 3 x 1
        g.take_time_step(0.0, 3)
        U = g.get_state()
        if normals == "exact":
            assert np.max(np.abs(U - U0)) < 1e-12
        I = g.integrals()
        assert np.allclose(I[:5], I0[:5], rtol=1e-12, atol=1e-13)
        g.close()

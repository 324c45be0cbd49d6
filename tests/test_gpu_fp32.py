"""The FP32 form of the stage (Numerics.precision = "FP32"): traces, Riemann
fluxes, point fluxes and every contraction of the face and volume phases in
single precision on FP32 copies of the state and the tables; the residual
is widened to FP64 before M^-1 and the RK update, which keep the FP64
registers.  Checked against the FP64 oracle:

* residual: max-norm error relative to max |R| below 2e-5 (the residual is
  a difference of volume and face terms that are each ~10x larger, so
  FP32 rounding of the terms, ~6e-8 relative, shows at ~1e-6 of R);
* the 8^3 p=2 vortex at CFL 0.1, 100 RK4 steps: the L2 error against the
  exact solution within 1.05x of the FP64 run's, the state within 1e-5;
* a uniform state stays uniform to FP32 rounding (1e-6); mass and energy
  stay constant to 1e-6 relative (Legendre: the mean mode has no volume
  term; the face terms cancel to FP32 rounding of the lift);
* Numerics.subtract_mean (on by default): the stage subtracts the domain
  mean from the coefficients before rounding them and forms deviation
  fluxes F(Ur + U') - F(Ur); the residual of the state set by
  init_state_from_fcn (whose mean is the reference) is checked against the
  oracle with the mean subtracted and without, on every flux, both bases and
  the orders whose code differs (p = 1 and 2: the face-normal-frame
  solver; p = 3: one solver per direction; p = 5: the 12-value point
  form).
"""
import numpy as np
import pytest

import dg_amd
import oracle as O
from dg_amd import hip

pytestmark = pytest.mark.gpu

VORTEX = ("ExtrudedIsentropicVortex", (1.0, 1.0, 1.0, 1.0, 5.0))


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def gpu():
    assert hip.device_count() > 0, "GPU test needs a GPU"


@pytest.mark.parametrize("order,basis,flux", [(2, "LegendreHex", "Roe"), (3, "LagrangeHex", "HLLC"),
                                              (3, "LegendreHex", "LaxFriedrichs"), (1, "LegendreHex", "Roe"),
                                              (4, "LagrangeHex", "Roe"), (3, "LegendreHex", "Roe")])
def test_fp32_residual_against_the_oracle(gpu, order, basis, flux):
    ne = (4, 3, 5)
    with dg_amd.Solver(dg_amd.case_toml(order=order, basis=basis, flux=flux, nelem=ne, timestep=0.01,
                                        ic=VORTEX[0], ic_data=VORTEX[1], precision="FP32")) as g:
        assert g.info.precision == 32
        o = O.Solver(O.make_params(order=order, basis=basis, flux=flux, nelem=ne, lo=(-5, -5, -5), hi=(5, 5, 5),
                                   ic=VORTEX[0], ic_data=VORTEX[1]))
        o.init_state()
        U = o.get_state()
        Ud = hip.Buffer.from_array(U)
        Rd = hip.Buffer(U.size)
        g.get_residual(Ud.addr, Rd.addr)
        hip.synchronize()
        R = Rd.download(U.shape)
        e = relerr(R, o.residual(U))
        print(f"p={order} {basis} {flux}: FP32 residual rel. error {e:.2e}", flush=True)
        assert e < 2e-5


def test_fp32_vortex_8cube_p2_100_steps(gpu):
    res = {}
    for prec in (None, "FP32"):
        toml = dg_amd.case_toml(order=2, basis="LegendreHex", flux="Roe", nelem=(8, 8, 8), stepper="RK4",
                                cfl=0.1, end=100.0, ic=VORTEX[0], ic_data=VORTEX[1], precision=prec)
        with dg_amd.Solver(toml) as g:
            g.init_state_from_fcn()
            g.take_time_step(g.timestep, 100)
            res[prec] = (g.get_state(), g.get_error(), g.timestep)
    U64, e64, dt64 = res[None]
    U32, e32, dt32 = res["FP32"]
    assert dt32 == dt64
    d = relerr(U32, U64)
    print("L2 errors FP64", e64, "FP32", e32, f"state rel. diff {d:.2e}", flush=True)
    assert d < 1e-5
    assert np.all(e32[[0, 1, 2, 4]] <= 1.05 * e64[[0, 1, 2, 4]])


def test_fp32_free_stream_and_conservation(gpu):
    with dg_amd.Solver(dg_amd.case_toml(order=3, basis="LagrangeHex", flux="HLLC", nelem=(3, 2, 4), timestep=0.01,
                                        ic="Uniform", ic_data=(1.2, 0.3, -0.2, 0.1, 3.1), precision="FP32")) as g:
        g.init_state_from_fcn()
        U0 = g.get_state()
        g.take_time_step(0.0, 3)
        assert np.max(np.abs(g.get_state() - U0)) < 1e-6
    with dg_amd.Solver(dg_amd.case_toml(order=3, basis="LegendreHex", flux="Roe", nelem=(4, 4, 4), timestep=0.01,
                                        ic=VORTEX[0], ic_data=VORTEX[1], precision="FP32")) as g:
        g.init_state_from_fcn()
        I0 = g.integrals()
        g.take_time_step(0.0, 10)
        I = g.integrals()
        assert abs(I[0] - I0[0]) < 1e-6 * abs(I0[0])
        assert abs(I[4] - I0[4]) < 1e-6 * abs(I0[4])


@pytest.mark.parametrize("order,basis,flux", [(1, "LegendreHex", "Roe"), (2, "LagrangeHex", "HLLC"),
                                              (3, "LegendreHex", "Roe"), (3, "LagrangeHex", "LaxFriedrichs"),
                                              (3, "LegendreHex", "HLLC"), (5, "LegendreHex", "Roe")])
def test_fp32_mean_subtracted_residual(gpu, order, basis, flux):
    ne = (4, 3, 5)
    errs = {}
    for sub in (False, True):
        with dg_amd.Solver(dg_amd.case_toml(order=order, basis=basis, flux=flux, nelem=ne, timestep=0.01,
                                            ic=VORTEX[0], ic_data=VORTEX[1], precision="FP32",
                                            subtract_mean=sub)) as g:
            g.init_state_from_fcn()
            U = g.get_state()
            o = O.Solver(O.make_params(order=order, basis=basis, flux=flux, nelem=ne, lo=(-5, -5, -5),
                                       hi=(5, 5, 5), ic=VORTEX[0], ic_data=VORTEX[1]))
            Ud = hip.Buffer.from_array(U)
            Rd = hip.Buffer(U.size)
            g.get_residual(Ud.addr, Rd.addr)
            hip.synchronize()
            errs[sub] = relerr(Rd.download(U.shape), o.residual(U))
    print(f"p={order} {basis} {flux}: FP32 residual rel. error {errs[False]:.2e} plain, "
          f"{errs[True]:.2e} mean subtracted", flush=True)
    assert errs[True] < 2e-5 and errs[False] < 2e-5
    assert errs[True] < errs[False]

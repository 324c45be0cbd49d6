"""Stepper.compensated (the solution update U0 + increment carried as the
FP64 sum and its TwoSum rounding error) on the GPU.

* Many small steps: 3000 RK4 steps of dt = 1e-4 on a 3^3 p = 1 box, where
  each increment is ~1e-4 of the state and the plain update rounds away
  its low bits every step.  The reference is RK4 with the solution
  accumulated in long double (numpy longdouble, 64-bit mantissa) around the
  library's own residual and M^-1 (dg_residual, dg_apply_inv_mass: the
  arithmetic of the fused stage, so that only the accumulation differs;
  the residual itself is pinned to the oracle elsewhere): the stage states
  are that solution rounded to FP64 plus alpha k, the increments summed in
  FP64 as the kernel does.  Distances in units of the spacing of the
  largest state value: the compensated state stays within 4 of it; the
  plain state drifts by the accumulated rounding, at least 4x further.
* Ordinary steps: 5 compensated RK4 and FE steps against the oracle at the
  plain bar (1e-11), and the FP32 form's default (compensated) run against
  its plain run (the compensation changes the state at FP64 rounding only).
"""
import numpy as np
import pytest

import dg_amd
import oracle as O

pytestmark = pytest.mark.gpu

VORTEX = ("ExtrudedIsentropicVortex", (1.0, 0.5, 0.3, 1.0, 3.0))


def toml(ne, stepper="RK4", dt=1e-4, comp=None, prec=None, order=1):
    return dg_amd.case_toml(order=order, basis="LegendreHex", flux="Roe", nelem=ne, stepper=stepper, timestep=dt,
                            ic=VORTEX[0], ic_data=VORTEX[1], compensated=comp, precision=prec)


def ld_reference(g, U0, dt, nsteps):
    """RK4 with the solution kept in long double (see the module text)."""
    from dg_amd import hip
    Ub, Rb, Kb = hip.Buffer(U0.size), hip.Buffer(U0.size), hip.Buffer(U0.size)

    def k(u):
        Ub.upload(u)
        g.get_residual(Ub.addr, Rb.addr)
        g.mult_inv_mass_matrix(dt, Rb.addr, Kb.addr)
        hip.synchronize()
        return Kb.download(U0.shape)

    U = U0.astype(np.longdouble)
    for _ in range(nsteps):
        u = U.astype(np.float64)
        k1 = k(u)
        k2 = k(u + 0.5 * k1)
        k3 = k(u + 0.5 * k2)
        k4 = k(u + k3)
        acc = k1 + 2.0 * k2 + 2.0 * k3
        U += ((acc + k4) * (1.0 / 6.0)).astype(np.longdouble)
    return U


def test_compensated_update_tracks_long_double_accumulation():
    ne, dt, n = (3, 3, 3), 1e-4, 3000
    res = {}
    for comp in (False, True):
        with dg_amd.Solver(toml(ne, dt=dt, comp=comp)) as g:
            g.init_state_from_fcn()
            U0 = g.get_state()
            g.take_time_step(dt, n)
            res[comp] = g.get_state()
    with dg_amd.Solver(toml(ne, dt=dt, comp=False)) as g:
        ref = ld_reference(g, U0, dt, n)
    ulp = float(np.spacing(np.max(np.abs(ref.astype(np.float64)))))
    err = {c: float(np.max(np.abs(res[c].astype(np.longdouble) - ref)) / ulp) for c in res}
    moved = float(np.max(np.abs(res[True] - U0)))
    print(f"{n} RK4 steps of dt {dt}: state moved by {moved:.3e}; distance to the long-double "
          f"accumulation in units of the largest value's spacing: plain {err[False]:.1f}, "
          f"compensated {err[True]:.2f}", flush=True)
    assert moved > 1e-3
    assert err[True] <= 4.0
    assert err[False] >= 4.0 * max(err[True], 1.0)


@pytest.mark.parametrize("stepper", ["RK4", "FE"])
def test_compensated_steps_against_the_oracle(stepper):
    ne, dt = (4, 3, 2), 0.01
    with dg_amd.Solver(toml(ne, stepper, dt, comp=True, order=2)) as g:
        g.init_state_from_fcn()
        U0 = g.get_state()
        g.take_time_step(dt, 5)
        U = g.get_state()
    o = O.Solver(O.make_params(order=2, basis="LegendreHex", flux="Roe", nelem=ne, lo=(-5, -5, -5), hi=(5, 5, 5),
                               ic=VORTEX[0], ic_data=VORTEX[1]))
    o.set_state(U0)
    o.step(stepper, dt, 5)
    ref = o.get_state()
    assert float(np.max(np.abs(U - ref)) / np.max(np.abs(ref))) < 1e-11


def test_fp32_form_defaults_to_the_compensated_update():
    ne, dt = (6, 5, 4), 0.01
    out = {}
    for comp in (None, False):
        with dg_amd.Solver(toml(ne, dt=dt, comp=comp, prec="FP32", order=3)) as g:
            g.init_state_from_fcn()
            g.take_time_step(dt, 20)
            out[comp] = g.get_state()
    d = float(np.max(np.abs(out[None] - out[False])) / np.max(np.abs(out[False])))
    print(f"FP32 form, compensated vs plain update after 20 steps: {d:.2e}", flush=True)
    # the two runs differ (the default compensates), by FP64 rounding carried
    # through 20 FP32 residual evaluations: below the FP32 form's own error
    assert 0.0 < d < 1e-6

"""Residuals of states with strong jumps at every face, against the oracle.

The smooth inputs of the other parity tests keep the Riemann solvers in
their subsonic, weak-jump branches.  Here each element holds a constant
state drawn from the left and right states of Toro's five Riemann problems
(E. F. Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics,
3rd ed., Table 4.1): a shock tube, the "123" problem (two strong
rarefactions, near-vacuum between them), pressure ratios of 1e5 (strong
shocks), two colliding shocks, and flow at Mach ~ 16 in which every wave
moves one way (HLLC's supersonic branches, Roe's entropy fix).  Each state's
velocity points along a seeded random direction, so every face of the box
sees a different jump.  A constant state is the coefficient 1 of the
constant function in the basis (Legendre: mode 0 only; Lagrange: every
node).

Bars: the FP64 residual within 1e-12 of max |R| of the oracle's (the
parity bar of the smooth cases); the FP32 form (domain mean subtracted, the
default) within 2e-5, its bar for smooth states.  The 2D quadrilateral path
(its own Riemann solvers) is checked the same way against oracle/dg2d.py."""
import numpy as np
import pytest

import dg_amd
import dg2d
import oracle as O
from dg_amd import hip

pytestmark = pytest.mark.gpu

GAMMA = 1.4
# (rho, u, p) left and right states of Toro's tests 1 - 5
TORO = [((1.0, 0.0, 1.0), (0.125, 0.0, 0.1)),
        ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4)),
        ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01)),
        ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950)),
        ((1.0, -19.59745, 1000.0), (1.0, -19.59745, 0.01))]


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def jump_state(order, basis, ne, seed):
    """[E][(p+1)^3][5] coefficients: element e holds a Toro state (test and
    side drawn per element), velocity along a random unit vector."""
    rng = np.random.default_rng(seed)
    E, nb = int(np.prod(ne)), (order + 1) ** 3
    U = np.zeros((E, nb, 5))
    for e in range(E):
        t, side = rng.integers(5), rng.integers(2)
        rho, u, p = TORO[t][side]
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        m = rho * u * n
        s = np.array([rho, m[0], m[1], m[2], p / (GAMMA - 1.0) + 0.5 * rho * u * u])
        if basis == "LegendreHex":
            U[e, 0] = s
        else:
            U[e, :] = s
    return U


def gpu_residual(order, basis, flux, ne, U, precision=None):
    toml = dg_amd.case_toml(order=order, basis=basis, flux=flux, nelem=ne, timestep=0.01, gamma=GAMMA,
                            precision=precision)
    with dg_amd.Solver(toml) as g:
        g.set_state(U)                  # the FP32 form takes its reference state from it
        Ud = hip.Buffer.from_array(U)
        Rd = hip.Buffer(U.size)
        g.get_residual(Ud.addr, Rd.addr)
        hip.synchronize()
        return Rd.download(U.shape)


@pytest.mark.parametrize("order,basis", [(1, "LegendreHex"), (2, "LagrangeHex"), (3, "LegendreHex"),
                                         (4, "LagrangeHex")])
@pytest.mark.parametrize("flux", ["LaxFriedrichs", "HLLC", "Roe"])
def test_jump_residual_against_the_oracle(order, basis, flux):
    ne = (4, 3, 5)
    U = jump_state(order, basis, ne, seed=11 * order + len(flux))
    o = O.Solver(O.make_params(order=order, basis=basis, flux=flux, nelem=ne, lo=(-5, -5, -5), hi=(5, 5, 5),
                               gamma=GAMMA))
    ref = o.residual(U)
    assert np.all(np.isfinite(ref))
    e64 = relerr(gpu_residual(order, basis, flux, ne, U), ref)
    e32 = relerr(gpu_residual(order, basis, flux, ne, U, "FP32"), ref)
    print(f"p={order} {basis} {flux}, Toro-state jumps: residual rel. error FP64 {e64:.2e}, FP32 {e32:.2e}",
          flush=True)
    assert e64 < 1e-12
    assert e32 < 2e-5


def jump_state_2d(order, basis, ne, seed):
    rng = np.random.default_rng(seed)
    E, nb = int(np.prod(ne)), (order + 1) ** 2
    U = np.zeros((E, nb, 4))
    for e in range(E):
        t, side = rng.integers(5), rng.integers(2)
        rho, u, p = TORO[t][side]
        a = rng.uniform(0.0, 2.0 * np.pi)
        s = np.array([rho, rho * u * np.cos(a), rho * u * np.sin(a), p / (GAMMA - 1.0) + 0.5 * rho * u * u])
        if basis == "LegendreQuad":
            U[e, 0] = s
        else:
            U[e, :] = s
    return U


@pytest.mark.parametrize("order,basis", [(1, "LagrangeQuad"), (2, "LegendreQuad"), (3, "LagrangeQuad")])
@pytest.mark.parametrize("flux", ["LaxFriedrichs", "HLLC", "Roe"])
def test_jump_residual_2d_against_the_oracle(order, basis, flux):
    ne, lo, hi = (5, 4), (-5.0, -5.0), (5.0, 5.0)
    U = jump_state_2d(order, basis, ne, seed=7 * order + len(flux))
    o = dg2d.Box2D(order, basis, flux, ne, lo, hi, gamma=GAMMA)
    ref = o.residual(U)
    assert np.all(np.isfinite(ref))
    toml = dg_amd.case_toml(order=order, basis=basis, flux=flux, nelem=ne, lower=lo, upper=hi, timestep=0.01,
                            gamma=GAMMA, ic="IsentropicVortex", ic_data=(1.0, 1.0, 1.0, 1.0, 5.0))
    with dg_amd.Solver(toml) as g:
        Ud = hip.Buffer.from_array(U)
        Rd = hip.Buffer(U.size)
        g.get_residual(Ud.addr, Rd.addr)
        hip.synchronize()
        e = relerr(Rd.download(U.shape), ref)
    print(f"2D p={order} {basis} {flux}, Toro-state jumps: residual rel. error {e:.2e}", flush=True)
    assert e < 1e-12

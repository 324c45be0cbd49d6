"""The bf16 matrix-core forms of the stage (Numerics.precision = "BF16" and
"BF16x3"): the FP32 form with the x-y part of the weak-gradient projection
as one v_mfma_f32_16x16x16_bf16 GEMM per element (FP32 accumulation).  BF16
rounds both operands (the projection matrix and the z-contracted fluxes) to
bf16; BF16x3 splits each into a bf16 value and its bf16 rounding error and
sums three products, which carries ~16 significant bits.

Bars, stated from the rounding each form makes (relative to max |R| of the
FP64 oracle's residual): the projection terms are up to ~10x the residual,
so bf16 rounding (2^-9 per operand) shows at up to ~3e-2 of max |R| and the
split form's (~2^-17) at up to ~5e-4; the FP32 form's bar is 2e-5.  The
forms' own ordering is checked too: BF16x3 is at least 20x closer to the
oracle than BF16.  With the domain mean subtracted before the rounding
(Numerics.subtract_mean, the default; the reference set from the state by
set_state) the projection operands are deviation fluxes and every form is
closer to the oracle than without.  Mass is conserved exactly by the projection (the
Legendre mean mode has a zero gradient row: its projection row is 0 in
bf16 as well), so the total mass stays constant to FP32 rounding of the
face terms.  Over 100 steps of the 8^3 p = 2 vortex, the L2 errors are
printed against the FP64 run's; BF16x3 stays within 1.05x of it.
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


def residual(order, basis, flux, ne, prec, U, sub=False):
    with dg_amd.Solver(dg_amd.case_toml(order=order, basis=basis, flux=flux, nelem=ne, timestep=0.01,
                                        ic=VORTEX[0], ic_data=VORTEX[1], precision=prec,
                                        subtract_mean=sub)) as g:
        if prec in ("BF16", "BF16x3"):
            assert g.info.precision == 16
        g.set_state(U)
        Ud = hip.Buffer.from_array(U)
        Rd = hip.Buffer(U.size)
        g.get_residual(Ud.addr, Rd.addr)
        hip.synchronize()
        return Rd.download(U.shape)


@pytest.mark.parametrize("order,basis,flux", [(1, "LegendreHex", "Roe"), (2, "LagrangeHex", "HLLC"),
                                              (3, "LegendreHex", "Roe"), (3, "LagrangeHex", "LaxFriedrichs"),
                                              (4, "LegendreHex", "Roe"), (5, "LegendreHex", "HLLC")])
def test_bf16_residual_against_the_oracle(order, basis, flux):
    ne = (4, 3, 5)
    o = O.Solver(O.make_params(order=order, basis=basis, flux=flux, nelem=ne, lo=(-5, -5, -5), hi=(5, 5, 5),
                               ic=VORTEX[0], ic_data=VORTEX[1]))
    o.init_state()
    U = o.get_state()
    ref = o.residual(U)
    e = {p: relerr(residual(order, basis, flux, ne, p, U), ref) for p in ("FP32", "BF16x3", "BF16")}
    em = {p: relerr(residual(order, basis, flux, ne, p, U, sub=True), ref) for p in ("FP32", "BF16x3", "BF16")}
    print(f"p={order} {basis} {flux}: residual rel. error FP32 {e['FP32']:.2e}, BF16x3 {e['BF16x3']:.2e}, "
          f"BF16 {e['BF16']:.2e}; mean subtracted: FP32 {em['FP32']:.2e}, BF16x3 {em['BF16x3']:.2e}, "
          f"BF16 {em['BF16']:.2e}", flush=True)
    assert e["FP32"] < 2e-5
    assert e["BF16x3"] < 5e-4
    assert e["BF16"] < 3e-2
    assert e["BF16x3"] * 20 < e["BF16"]
    for p in e:
        assert em[p] < e[p], p


@pytest.mark.parametrize("prec", ["BF16", "BF16x3"])
def test_bf16_vortex_8cube_p2_100_steps(prec):
    res = {}
    for pr in (None, prec):
        toml = dg_amd.case_toml(order=2, basis="LegendreHex", flux="Roe", nelem=(8, 8, 8), stepper="RK4",
                                cfl=0.1, end=100.0, ic=VORTEX[0], ic_data=VORTEX[1], precision=pr)
        with dg_amd.Solver(toml) as g:
            g.init_state_from_fcn()
            I0 = g.integrals()
            g.take_time_step(g.timestep, 100)
            res[pr] = (g.get_state(), np.array(g.get_error()), I0, g.integrals())
    U64, e64, _, _ = res[None]
    U, e, I0, I1 = res[prec]
    d = relerr(U, U64)
    ratio = e[[0, 1, 2, 4]] / e64[[0, 1, 2, 4]]
    print(f"{prec}: 8^3 p=2 vortex after 100 steps, L2 error ratio vs FP64 {ratio}, state rel. diff {d:.2e}",
          flush=True)
    assert np.all(np.isfinite(U))
    assert abs(I1[0] - I0[0]) < 1e-6 * abs(I0[0])
    if prec == "BF16x3":
        assert np.all(ratio < 1.05), ratio
        assert d < 1e-3
    else:
        assert np.all(ratio < 2.0), ratio

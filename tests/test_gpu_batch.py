"""Batches of independent cases (dg_batch_*): uneven sizes, orders, fluxes,
vortex strengths and free-stream speeds, one case in the FP32 form, time
steps given or CFL-derived.  After each replay of the captured step graph
(with the program's own eager work between replays) every case holds the
time and, to 1e-14 of max |U|, the state that its own dg_step produces
(the batched kernel inlines the same stage code, but the compiler may fuse
a multiply and an add differently there; the differences seen are in the
last bits of round-off-sized entries such as rho w ~ 1e-35 of the
z-extruded vortex); one case is also checked against the oracle."""
import ctypes as C

import numpy as np
import pytest

import dg_amd
import oracle as O
from dg_amd import hip

pytestmark = pytest.mark.gpu

CASES = [
    # order, basis, flux, nelem, (rho, u, v, p, beta), timestep (None: CFL 0.1), precision
    (1, "LegendreHex", "Roe", (5, 4, 3), (1.0, 0.5, 0.3, 1.0, 2.0), None, None),
    (2, "LagrangeHex", "HLLC", (3, 3, 3), (1.0, 1.0, 1.0, 1.0, 5.0, 1.3, -0.7), 0.01, None),
    (3, "LegendreHex", "LaxFriedrichs", (4, 2, 3), (1.0, 0.2, -0.4, 1.0, 3.0), None, None),
    (4, "LegendreHex", "Roe", (2, 3, 2), (1.0, 0.7, 0.1, 1.0, 1.0), 0.005, None),
    (2, "LegendreHex", "Roe", (6, 6, 6), (1.0, -0.3, 0.6, 1.0, 4.0), None, None),
    (3, "LegendreHex", "Roe", (4, 4, 4), (1.0, 1.0, 1.0, 1.0, 5.0), None, "FP32"),
    (1, "LagrangeHex", "HLLC", (8, 8, 8), (1.0, 0.4, 0.4, 1.0, 2.5), 0.02, None),
    (3, "LegendreHex", "Roe", (3, 5, 4), (1.0, 0.9, -0.2, 1.0, 1.5), None, None),
]


def toml(c):
    order, basis, flux, ne, data, dt, prec = c
    return dg_amd.case_toml(order=order, basis=basis, flux=flux, nelem=ne, timestep=dt,
                            cfl=None if dt else 0.1, ic="ExtrudedIsentropicVortex", ic_data=data, precision=prec)


@pytest.fixture(scope="module")
def gpu():
    assert hip.device_count() > 0, "GPU test needs a GPU"


def test_batch_replays_match_each_case_alone(gpu):
    batch_s = [dg_amd.Solver(toml(c)) for c in CASES]
    solo_s = [dg_amd.Solver(toml(c)) for c in CASES]
    for s in batch_s + solo_s:
        s.init_state_from_fcn()
    with dg_amd.Batch(batch_s) as b:
        for rep in range(3):
            b.step(2)
            for s in solo_s:
                s.take_time_step(0.0, 2)
            for i, (sb, so) in enumerate(zip(batch_s, solo_s)):
                assert sb.time == so.time, (rep, i)
                Ub, Uo = sb.get_state(), so.get_state()
                d = float(np.max(np.abs(Ub - Uo)) / np.max(np.abs(Uo)))
                assert d < 1e-14, (rep, i, d)
            # eager work on the batch's handles between replays
            errs = [s.get_error() for s in batch_s]
            assert all(np.all(np.isfinite(e)) for e in errs)
            print(f"replay {rep}: ok, rho errors {[f'{e[0]:.2e}' for e in errs]}", flush=True)
    # the 3^3 p = 2 HLLC case against the oracle after its 6 batched steps
    c = CASES[1]
    o = O.Solver(O.make_params(order=c[0], basis=c[1], flux=c[2], nelem=c[3], lo=(-5, -5, -5), hi=(5, 5, 5),
                               ic="ExtrudedIsentropicVortex", ic_data=c[4]))
    o.init_state()
    o.step("RK4", c[5], 6)
    R = o.get_state()
    assert np.max(np.abs(batch_s[1].get_state() - R)) / np.max(np.abs(R)) < 1e-11
    for s in batch_s + solo_s:
        s.close()


def test_batch_input_checks(gpu):
    def create(handles):
        arr = (C.c_void_p * len(handles))(*[h._h.value for h in handles])
        h = C.c_void_p()
        rc = dg_amd.lib().dg_batch_create(len(handles), arr, C.byref(h))
        msg = dg_amd.lib().dg_batch_last_error(None).decode()
        if rc == 0:
            dg_amd.lib().dg_batch_destroy(h)
        return rc, msg
    a = dg_amd.Solver(toml(CASES[1]))
    warped = dg_amd.Solver(dg_amd.case_toml(order=2, nelem=(3, 3, 3), timestep=0.01, warp=0.1))
    fe = dg_amd.Solver(dg_amd.case_toml(order=2, nelem=(3, 3, 3), timestep=0.01, stepper="FE"))
    assert create([a, a]) == (3, "dg_batch_create: a case appears twice")
    rc, msg = create([a, warped])
    assert rc == 4 and "affine periodic box" in msg, msg
    rc, msg = create([fe])
    assert rc == 4 and "RK4" in msg, msg
    nodt = dg_amd.Solver(dg_amd.case_toml(order=2, nelem=(3, 3, 3), cfl=0.1))
    with dg_amd.Batch([nodt]) as b:
        with pytest.raises(dg_amd.DGError, match="no time step"):
            b.step(1)
    for s in (a, warped, fe, nodt):
        s.close()


def test_batch_tuning_keeps_results(gpu):
    """A call with nsteps >= 6 times, on its first six steps, the candidates
    of every group (two team sizes x bricks of at most 8^3 or 4^3 elements)
    and keeps the fastest; the cases still end where their own dg_step takes
    them, to round-off (a team size of its own compiles to its own
    multiply-add fusion: ~1e-16 relative in FP64, ~1e-8 for the FP32-form
    case; the traversal order changes nothing), also on replays of the
    re-captured step, whose groups run as concurrent graph branches."""
    cases = [CASES[0], CASES[2], CASES[3], CASES[5], CASES[6]]
    batch_s = [dg_amd.Solver(toml(c)) for c in cases]
    solo_s = [dg_amd.Solver(toml(c)) for c in cases]
    for s in batch_s + solo_s:
        s.init_state_from_fcn()
    with dg_amd.Batch(batch_s) as b:
        assert all(t[2] == 0.0 for t in b.tuning())
        b.step(6)
        tun = b.tuning()
        print("tuning (order, team, ms default, ms alternate):", tun, flush=True)
        assert len(tun) == len(cases)
        for order, team, t0, t1 in tun:
            assert t0 > 0.0 and t1 > 0.0
            assert team in (64, 128, 256, 512)
        tiles = b.tiles()
        print("tiles (brick cap, cases, elements, ms per candidate):", tiles, flush=True)
        assert sum(t[1] for t in tiles) == len(cases)
        for (cap, nc, ne, ms), (order, team, t0, t1) in zip(tiles, tun):
            assert cap in (3, 2) and len(ms) == 4 and min(ms) > 0.0
            ci = 0 if cap == 3 else 1
            assert min(ms) == min(ms[2 * ci], ms[2 * ci + 1])      # the chosen cap holds the fastest candidate
            assert (t0, t1) == (ms[2 * ci], ms[2 * ci + 1])
        for s in solo_s:
            s.take_time_step(0.0, 6)
        tol = [1e-14 if c[6] is None else 2e-6 for c in cases]
        for sb, so, t in zip(batch_s, solo_s, tol):
            assert sb.time == so.time
            Ub, Uo = sb.get_state(), so.get_state()
            assert np.max(np.abs(Ub - Uo)) / np.max(np.abs(Uo)) < t
        b.step(2)                               # replays of the re-captured step
        for s in solo_s:
            s.take_time_step(0.0, 2)
        for sb, so, t in zip(batch_s, solo_s, tol):
            Ub, Uo = sb.get_state(), so.get_state()
            assert np.max(np.abs(Ub - Uo)) / np.max(np.abs(Uo)) < t
    for s in batch_s + solo_s:
        s.close()

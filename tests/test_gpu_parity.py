"""GPU parity of the HIP product path (through the C ABI) with the CPU
oracle on the same seeded inputs, plus size-independent properties.

Bar: FP64 results within 1e-12 relative (max-norm over the array, relative
to the max-norm of the oracle result) unless stated otherwise."""
import json
import os

import numpy as np
import pytest

import dg_amd
import oracle as O
from dg_amd import hip

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def gpu():
    assert hip.device_count() > 0, "GPU test needs a GPU"
    return hip


def residual_of(g, U):
    Ud = hip.Buffer.from_array(U)
    Rd = hip.Buffer(U.size)
    g.get_residual(Ud.addr, Rd.addr)
    hip.synchronize()
    return Rd.download(U.shape)


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [
    # order, basis, flux, nelem, lo, hi, ic, data
    (1, "LagrangeHex", "LaxFriedrichs", (3, 2, 4), (-5, -5, -5), (5, 5, 5), "ExtrudedIsentropicVortex", (1, 1, 1, 1, 5)),
    (2, "LegendreHex", "Roe", (4, 4, 4), (-5, -5, -5), (5, 5, 5), "ExtrudedIsentropicVortex", (1, 1, 1, 1, 5)),
    (2, "LagrangeHex", "HLLC", (3, 4, 5), (0, -1, 2), (2, 1, 3.5), "GaussianBump",
     (0.2, 0.3, 0.6, -0.48, 0.64, 1.0, 1.0, 0.0, 2.7)),
    (3, "LagrangeHex", "HLLC", (4, 4, 4), (0, 0, 0), (1, 1, 1), "IsentropicVortex", (1, 1, 1, 1, 5)),
    (3, "LegendreHex", "LaxFriedrichs", (2, 3, 2), (-5, -5, -5), (5, 5, 5), "ExtrudedIsentropicVortex", (1, 1, 1, 1, 5)),
    (3, "LegendreHex", "Roe", (4, 3, 2), (0, 0, 0), (6.283185307179586,) * 3, "TaylorGreen", (0.1,)),
    (4, "LegendreHex", "Roe", (2, 2, 3), (-5, -5, -5), (5, 5, 5), "ExtrudedIsentropicVortex", (1, 1, 1, 1, 5)),
    (4, "LagrangeHex", "HLLC", (2, 3, 2), (-5, -5, -5), (5, 5, 5), "ExtrudedIsentropicVortex", (1, 1, 1, 1, 5)),
    (5, "LagrangeHex", "Roe", (2, 2, 2), (-5, -5, -5), (5, 5, 5), "ExtrudedIsentropicVortex", (1, 1, 1, 1, 5)),
    # the headline configuration's discretisation (p=3 LegendreHex, Roe, the
    # z-extruded vortex on [-5,5]^3) on a small box
    (3, "LegendreHex", "Roe", (4, 4, 4), (-5, -5, -5), (5, 5, 5), "ExtrudedIsentropicVortex", (1, 1, 1, 1, 5)),
]


def make_pair(case, stepper="RK4", timestep=0.01, gamma=1.4):
    order, basis, flux, ne, lo, hi, ic, data = case
    toml = dg_amd.case_toml(order=order, basis=basis, flux=flux, nelem=ne, lower=lo, upper=hi,
                            stepper=stepper, timestep=timestep, end=1.0, gamma=gamma, ic=ic, ic_data=data)
    g = dg_amd.Solver(toml)
    o = O.Solver(O.make_params(order=order, basis=basis, flux=flux, nelem=ne, lo=lo, hi=hi,
                               gamma=gamma, R=1.0, ic=ic, ic_data=data))
    return g, o


def case_id(c):
    return f"p{c[0]}-{c[1][:4]}-{c[2]}-{c[6][:8]}"


@pytest.mark.parametrize("case", CASES, ids=case_id)
def test_init_state_projection(case):
    g, o = make_pair(case)
    print(f"case {case_id(case)}", flush=True)
    g.init_state_from_fcn()
    o.init_state()
    assert relerr(g.get_state(), o.get_state()) < TOL


@pytest.mark.parametrize("case", CASES, ids=case_id)
def test_residual(case, gpu):
    g, o = make_pair(case)
    print(f"case {case_id(case)}", flush=True)
    o.init_state()
    U = o.get_state()
    # perturb so that every basis coefficient is exercised
    rng = np.random.default_rng(1234)
    U = U * (1.0 + 1e-3 * rng.standard_normal(U.shape))
    Rg = residual_of(g, U)
    Ro = o.residual(U)
    assert np.all(np.isfinite(Rg))
    assert relerr(Rg, Ro) < TOL, relerr(Rg, Ro)
    # deterministic: bitwise identical on a second evaluation
    assert np.array_equal(residual_of(g, U), Rg)


@pytest.mark.parametrize("case", [CASES[1], CASES[3], CASES[7]], ids=case_id)
def test_inv_mass(case, gpu):
    g, o = make_pair(case)
    rng = np.random.default_rng(7)
    R = rng.standard_normal(g.shape)
    Rd = hip.Buffer.from_array(R)
    dUd = hip.Buffer(R.size)
    g.mult_inv_mass_matrix(0.37, Rd.addr, dUd.addr)
    hip.synchronize()
    assert relerr(dUd.download(R.shape), o.inv_mass(0.37, R)) < TOL
    # in place
    g.mult_inv_mass_matrix(0.37, Rd.addr, Rd.addr)
    hip.synchronize()
    assert relerr(Rd.download(R.shape), o.inv_mass(0.37, R)) < TOL


@pytest.mark.parametrize("stepper", ["RK4", "FE"])
@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[5], CASES[6], CASES[9]], ids=case_id)
def test_time_steps(case, stepper):
    dt = 0.002 if case[6] != "TaylorGreen" else 0.01
    g, o = make_pair(case, stepper=stepper, timestep=dt)
    print(f"case {case_id(case)} {stepper}", flush=True)
    o.init_state()
    U0 = o.get_state()
    g.set_state(U0)
    nsteps = 5
    g.take_time_step(0.0, nsteps)
    o.step(stepper, dt, nsteps)
    assert relerr(g.get_state(), o.get_state()) < 1e-11
    assert abs(g.time - o.time) < 1e-15


@pytest.mark.parametrize("case", [CASES[1], CASES[3], CASES[6]], ids=case_id)
def test_l2_error_and_wavespeed(case):
    g, o = make_pair(case)
    o.init_state()
    U = o.get_state()
    g.set_state(U)
    for t in (0.0, 0.3):
        eg, eo = g.get_error(t), o.l2_error(t)
        assert np.allclose(eg, eo, rtol=1e-11, atol=1e-15), (eg, eo)
    assert abs(g.max_wavespeed() - o.max_wavespeed()) < 1e-13 * o.max_wavespeed()


def test_reference_3d_vortex_regression(golden_dir):
    """examples/euler/3D/isentropic_vortex/regression_data.txt:11-17 on the
    regenerated periodic 4^3 unit cube: p=3, LagrangeHex, HLLC, RK4,
    dt=1e-3, end=1.0 (1000 steps), through dg_solve."""
    gd = json.load(open(os.path.join(golden_dir, "vortex3d_regression.json")))
    toml = dg_amd.case_toml(order=3, basis="LagrangeHex", flux="HLLC", nelem=(4, 4, 4), lower=(0, 0, 0),
                            upper=(1, 1, 1), stepper="RK4", timestep=gd["timestep"], end=gd["end"],
                            ic="IsentropicVortex", ic_data=(1.0, 1.0, 1.0, 1.0, 5.0))
    g = dg_amd.Solver(toml)
    g.init_state_from_fcn()
    assert g.num_iterations == 1000
    g.solve()
    e = g.get_error()
    ref = np.array([float(v) for v in gd["l2_errors"]])
    print("L2 errors", e, flush=True)
    for v in (0, 1, 3, 4):
        assert abs(e[v] - ref[v]) < 1e-10 * ref[v], (v, e[v], ref[v])
    assert e[2] < 1e-14


@pytest.mark.parametrize("flux", ["LaxFriedrichs", "HLLC", "Roe"])
def test_free_stream_and_conservation(flux, gpu):
    # free stream: R = 0 to round-off
    toml = dg_amd.case_toml(order=3, basis="LegendreHex", flux=flux, nelem=(6, 5, 4), lower=(-1, 0, 0),
                            upper=(2, 1, 3), timestep=0.01, ic="Uniform", ic_data=(1.2, 0.3, -0.2, 0.5, 3.1))
    g = dg_amd.Solver(toml)
    g.init_state_from_fcn()
    R = residual_of(g, g.get_state())
    assert float(np.abs(R).max()) < 1e-13
    # conservation: the element-mean (Legendre mode 0) residual sums to zero
    toml = dg_amd.case_toml(order=3, basis="LegendreHex", flux=flux, nelem=(6, 5, 4), timestep=0.01)
    g = dg_amd.Solver(toml)
    g.init_state_from_fcn()
    R = residual_of(g, g.get_state())
    tot = R[:, 0, :].sum(axis=0)
    assert np.all(np.abs(tot) < 1e-13 * np.abs(R[:, 0, :]).sum()), tot


def test_state_layouts():
    g, _ = make_pair(CASES[2])
    rng = np.random.default_rng(3)
    U = rng.standard_normal(g.shape)
    g.set_state(U)
    assert np.array_equal(g.get_state(), U)
    L = g.get_state(dg_amd.DG_LAYOUT_LEFT)
    assert np.array_equal(L, np.transpose(U, (2, 1, 0)))
    g.set_state(np.ascontiguousarray(L), dg_amd.DG_LAYOUT_LEFT)
    assert np.array_equal(g.get_state(), U)


def test_cfl_timestep_and_conservation_of_integrals():
    toml = dg_amd.case_toml(order=2, basis="LegendreHex", flux="Roe", nelem=(8, 8, 8), cfl=0.1, end=0.5)
    g = dg_amd.Solver(toml)
    g.init_state_from_fcn()
    o = O.Solver(O.make_params(order=2, basis="LegendreHex", flux="Roe", nelem=(8, 8, 8), lo=(-5, -5, -5),
                               hi=(5, 5, 5), ic="ExtrudedIsentropicVortex", ic_data=(1, 1, 1, 1, 5)))
    o.init_state()
    amax = o.max_wavespeed()
    assert abs(g.timestep - 0.1 * (10 / 8) / amax) < 1e-14
    I0 = g.integrals()
    g.take_time_step(0.0, 10)
    I1 = g.integrals()
    # mass, momentum and energy are conserved to round-off
    assert np.all(np.abs(I1[:5] - I0[:5]) < 1e-12 * np.abs(I0[:5]).max())


def test_dg_run_and_dg_post_on_the_reference_example(tmp_path, golden_dir):
    """dg_run / dg_post (counterparts of src/exec/main.cpp and
    src/post/main.cpp) on the case of examples/euler/3D/isentropic_vortex/
    input.toml with the HDF5 mesh replaced by the periodic-box keys."""
    import re
    import subprocess
    gd = json.load(open(os.path.join(golden_dir, "vortex3d_regression.json")))
    case = """[Mesh]
type = "PeriodicBox"
nelem = [4, 4, 4]
lower = [0.0, 0.0, 0.0]
upper = [1.0, 1.0, 1.0]

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
"""
    (tmp_path / "input.toml").write_text(case)
    bindir = os.path.join(os.path.dirname(dg_amd.PKG_ROOT), "dg-mpi_amd", "bin")
    r = subprocess.run([os.path.join(bindir, "dg_run")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Number of iterations = 1000" in r.stdout
    r = subprocess.run([os.path.join(bindir, "dg_post")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    errs = [float(v) for v in re.findall(r"norm_total_error\(\d\)=([-+0-9.eE]+)", r.stdout)]
    ref = [float(v) for v in gd["l2_errors"]]
    assert len(errs) == 5
    for v in (0, 1, 3, 4):
        assert abs(errs[v] - ref[v]) < 1e-10 * ref[v]

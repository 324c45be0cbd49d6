"""GPU parity of the 2D (quadrilateral) path, Physics.dim = 2: periodic
boxes of affine quadrilaterals, 4 state variables, faces on the (p+1)-point
segment rule (src/solver/helpers.cpp:225), against the 2D oracle
(oracle/dg2d.py) and against the reference's own 2D convergence data
(examples/isentropic_vortex/convergence.py:52-86).

Bar: 1e-12 relative (max-norm) for the residual; 1e-11 for results through
M^-1 and for five steps; the convergence data to 1e-6 relative or 1e-10
absolute (the data was computed on GPUs by the reference: rounding
differences of the solution after 1000 steps, ~1e-11, are a visible part of
the smallest errors)."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import dg_amd
import dg2d
from dg_amd import hip

pytestmark = pytest.mark.gpu

VORTEX = ("IsentropicVortex", (1.0, 1.0, 1.0, 1.0, 5.0))
CASES = [
    # order, basis, flux, nelem, lo, hi, (ic, data)
    (1, "LagrangeQuad", "LaxFriedrichs", (3, 4), (-5, -5), (5, 5), VORTEX),
    (2, "LegendreQuad", "Roe", (4, 3), (-5, -5), (5, 5), VORTEX),
    (3, "LagrangeQuad", "HLLC", (4, 4), (-5, -5), (5, 5), VORTEX),
    (3, "LegendreQuad", "LaxFriedrichs", (2, 5), (-4, -5), (6, 5), VORTEX),
    (4, "LegendreQuad", "HLLC", (3, 3), (-5, -5), (5, 5), VORTEX),
    (5, "LagrangeQuad", "Roe", (3, 2), (-5, -5), (5, 5), VORTEX),
]
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "vortex2d_convergence.json")))
BINDIR = os.path.join(os.path.dirname(dg_amd.PKG_ROOT), "dg-mpi_amd", "bin")


def case_id(c):
    return f"p{c[0]}-{c[1][:4]}-{c[2]}-{c[3][0]}x{c[3][1]}"


@pytest.fixture(scope="module")
def gpu():
    assert hip.device_count() > 0, "GPU test needs a GPU"


def toml(order, basis, flux, ne, lo, hi, ic, data, stepper="RK4", timestep=0.001, end=1.0):
    return dg_amd.case_toml(order=order, basis=basis, flux=flux, nelem=ne, lower=lo, upper=hi, stepper=stepper,
                            timestep=timestep, end=end, ic=ic, ic_data=data)


def make_pair(case, stepper="RK4", timestep=0.001):
    order, basis, flux, ne, lo, hi, (ic, data) = case
    g = dg_amd.Solver(toml(order, basis, flux, ne, lo, hi, ic, data, stepper, timestep))
    o = dg2d.Box2D(order, basis, flux, ne, lo, hi, ic=ic, ic_data=data)
    return g, o


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("case", CASES, ids=case_id)
def test_2d_projection_residual_inv_mass(gpu, case):
    print(f"case {case_id(case)}", flush=True)
    g, o = make_pair(case)
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
    g.close()


@pytest.mark.parametrize("stepper", ["RK4", "FE"])
@pytest.mark.parametrize("case", [CASES[0], CASES[1], CASES[4]], ids=case_id)
def test_2d_time_steps(gpu, case, stepper):
    print(f"case {case_id(case)} {stepper}", flush=True)
    g, o = make_pair(case, stepper=stepper, timestep=0.002)
    o.init_state()
    g.set_state(o.U)
    g.take_time_step(0.0, 5)
    o.step(stepper, 0.002, 5)
    assert relerr(g.get_state(), o.U) < 1e-11
    g.close()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_reference_2d_convergence_data(gpu, p):
    """All the sizes of the reference's 2D study, 1000 RK4 steps each."""
    errs = GOLD[f"err_p{p}"]
    bad = []
    for k, n in enumerate(GOLD["nelem_1d"][:len(errs)]):
        with dg_amd.Solver(toml(p, "LagrangeQuad", "HLLC", (n, n), (-5, -5), (5, 5), *VORTEX)) as g:
            g.init_state_from_fcn()
            g.solve()
            assert g.num_iterations == 1000
            e = g.get_error()[0]
        ref = float(errs[k])
        print(f"p={p} N={n} err={e:.15e} ref={ref:.15e} diff={abs(e - ref):.2e} rel={abs(e - ref) / ref:.2e}",
              flush=True)
        if abs(e - ref) > max(1e-6 * ref, 1e-10):
            bad.append((n, e, ref))
    assert not bad, bad


@pytest.mark.parametrize("flux", ["LaxFriedrichs", "HLLC", "Roe"])
def test_2d_free_stream_and_conservation(gpu, flux):
    with dg_amd.Solver(toml(3, "LagrangeQuad", flux, (3, 2), (0, 0), (3, 1), "Uniform", (1.2, 0.3, -0.2, 3.1),
                            timestep=0.01)) as g:
        g.init_state_from_fcn()
        U0 = g.get_state()
        g.take_time_step(0.0, 3)
        assert np.max(np.abs(g.get_state() - U0)) < 1e-13
    with dg_amd.Solver(toml(2, "LegendreQuad", flux, (4, 4), (-5, -5), (5, 5), *VORTEX, timestep=0.01)) as g:
        g.init_state_from_fcn()
        I0 = g.integrals()
        g.take_time_step(0.0, 10)
        I = g.integrals()
        assert np.allclose(I[:4], I0[:4], rtol=1e-13, atol=1e-12)
        assert I[5] == 0.0


def test_2d_gaussian_projection(gpu):
    """The reference's 2D "Gaussian" initial condition (a test function,
    physics/base/base.cpp:161-178), projected."""
    data = (1.0, 0.5, -0.25, 2.0)
    with dg_amd.Solver(toml(3, "LegendreQuad", "HLLC", (3, 4), (-1, -1), (1, 1), "Gaussian", data)) as g:
        o = dg2d.Box2D(3, "LegendreQuad", "HLLC", (3, 4), (-1, -1), (1, 1), ic="Gaussian", ic_data=data)
        g.init_state_from_fcn()
        o.init_state()
        assert relerr(g.get_state(), o.U) < 1e-11
        assert np.allclose(g.get_error(0.0)[:4], o.l2_error(0.0), rtol=1e-9, atol=1e-15)


def test_2d_solution_file_and_tools(gpu, tmp_path):
    """dg_run / dg_post on the reference's 2D example case with the mesh
    file replaced by the periodic box keys (16 x 16, p = 2): the density
    error of the reference's convergence data, a data.h5 with dimension 2."""
    h5 = pytest.importorskip("dg_amd.h5")
    if not os.path.exists(h5.H5_LIB_PATH):
        pytest.skip("libdg_amd_h5.so not built")
    (tmp_path / "input.toml").write_text("""[Mesh]
type = "PeriodicBox"
nelem = [16, 16]
lower = [-5.0, -5.0]
upper = [5.0, 5.0]

[Numerics]
order = 2
basis = "LagrangeQuad"

[Stepper]
start = 0.0
end = 1.0
timestep = 0.001
type = "RK4"

[Physics]
name = "Euler"
dim = 2
gamma = 1.4
GasConstant = 1.0
convective_flux_fcn = "HLLC"

[InitialCondition]
name = "IsentropicVortex"
data = [1.0, 1.0, 1.0, 1.0, 5.0]
""")
    r = subprocess.run([os.path.join(BINDIR, "dg_run")], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    hdr = h5.read_header(str(tmp_path / "data.h5"))
    assert (hdr["dim"], hdr["num_nodes_per_elem"], hdr["num_elems"], hdr["num_basis"], hdr["num_state_vars"]) == \
        (2, 4, 256, 9, 4)
    assert hdr["num_nodes"] == 17 * 17
    r = subprocess.run([os.path.join(BINDIR, "dg_post")], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    errs = [float(v) for v in re.findall(r"norm_total_error\(\d\)=([-+0-9.eE]+)", r.stdout)]
    assert len(errs) == 4
    ref = float(GOLD["err_p2"][GOLD["nelem_1d"].index(16)])
    assert abs(errs[0] - ref) < 1e-6 * ref

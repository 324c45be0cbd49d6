"""dg_closure_step_noalloc (the closure step without device allocation: its
scratch is a static device array of the library) against dg_closure_step on
the same inputs: bit-identical states and equal fallback counts, on two
calls in a row, the second with a smaller box; a box above the array's
32^3 elements is rejected with DG_ERR_ARGUMENT before any device work."""
import numpy as np
import pytest

import dg_amd
import mlp as MO
from dg_amd import closure

pytestmark = pytest.mark.gpu

VORTEX = ("ExtrudedIsentropicVortex", (1.0, 0.5, 0.3, 1.0, 3.0))


def net(seed):
    rng = np.random.default_rng(seed)
    p = np.concatenate([rng.normal(0, 0.2, int(np.prod(s))) for _, s in MO.SHAPES]).astype(np.float32)
    m = closure.MLP(0)
    m.set_params(p)
    m.set_scaling(rng.normal(0, 0.1, MO.NF).astype(np.float32), rng.uniform(0.5, 2.0, MO.NF).astype(np.float32),
                  np.full(MO.NO, 1e-2, np.float32))
    return m


def case(ne):
    return dg_amd.case_toml(order=1, basis="LegendreHex", flux="Roe", nelem=ne, timestep=0.01, ic=VORTEX[0],
                            ic_data=VORTEX[1])


@pytest.mark.parametrize("ne", [(6, 5, 4), (3, 3, 2)])
def test_noalloc_closure_step_matches_the_allocating_one(ne):
    print(f"closure noalloc vs dg_closure_step on {ne}", flush=True)
    m = net(11)
    with dg_amd.Solver(case(ne)) as g:
        g.init_state_from_fcn()
        U0 = g.get_state()
        # a bound that some elements exceed, so the fallback path is exercised
        n_ref = closure.step(g, m, 0.05, 2)
        U_ref = g.get_state()
        g.set_state(U0)
        g.set_time(0.0)
        n_na = closure.step_noalloc(g, m, 0.05, 2)
        U_na = g.get_state()
    m.close()
    assert n_na == n_ref
    assert np.array_equal(U_na, U_ref)
    assert not np.array_equal(U_ref, U0)


def test_noalloc_rejects_boxes_above_its_scratch():
    m = net(12)
    with dg_amd.Solver(case((33, 32, 32))) as g:
        g.init_state_from_fcn()
        with pytest.raises(dg_amd.DGError) as ei:
            closure.step_noalloc(g, m, 1e30, 1)
        assert ei.value.status == 3 and "32768" in str(ei.value)
    m.close()

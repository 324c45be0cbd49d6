"""The learned-closure kernels against the CPU oracle: the network's
inference pass (the matrix-core kernel of dg_mlp_forward and the closure
step: bf16 weights and layer inputs, FP32 accumulation; bar 1e-3 of the
largest output, since an activation that FP32 and float64 round to
neighbouring bf16 values moves an output by ~1e-4 of it) and its gradient
(oracle/mlp.py, float64 with the same bfloat16 weight rounding and FP32
activations; bar 1e-4 relative per parameter block), the coarse-state
features and the training targets (oracle/oracle.py residuals; FP32
outputs, bar 1e-6 relative), closure stepping with a network whose output
is zero against plain RK4 (1e-13), one corrected RK4 step of a random
network against the oracle, the fallback bound, and a trained network's
held-out gain."""
import ctypes as C

import numpy as np
import pytest

import dg_amd
import mlp as MO
import oracle as O
from dg_amd import closure, hip

pytestmark = pytest.mark.gpu

VORTEX = ("ExtrudedIsentropicVortex", (1.0, 0.5, 0.3, 1.0, 3.0))


@pytest.fixture(scope="module")
def gpu():
    assert hip.device_count() > 0, "GPU test needs a GPU"


def rand_net(seed=0):
    rng = np.random.default_rng(seed)
    p = np.concatenate([rng.normal(0, 0.2, int(np.prod(s))) for _, s in MO.SHAPES]).astype(np.float32)
    xm = rng.normal(0, 0.1, MO.NF).astype(np.float32)
    xs = rng.uniform(0.5, 2.0, MO.NF).astype(np.float32)
    ys = rng.uniform(0.5, 2.0, MO.NO).astype(np.float32)
    return p, xm, xs, ys


def test_mlp_forward_and_gradient_against_the_oracle(gpu):
    p, xm, xs, ys = rand_net()
    rng = np.random.default_rng(3)
    rows = 300
    X = rng.normal(0, 1, (rows, MO.NF)).astype(np.float32)
    Y = rng.normal(0, 1, (rows, MO.NO)).astype(np.float32)
    m = closure.MLP(0)
    m.set_params(p)
    m.set_scaling(xm, xs, ys)
    Xd, Yd = hip.Buffer.from_array(X, np.float32), hip.Buffer.from_array(Y, np.float32)
    out = m.forward(Xd, rows).download((rows, MO.NO))
    ref, _ = MO.forward(p, X, xm.astype(np.float64), xs.astype(np.float64), round_acts=True)
    ref = ref * ys
    e = np.max(np.abs(out - ref)) / np.max(np.abs(ref))
    # against FP32 activations: the bf16 rounding of the layer inputs shows
    ref32, _ = MO.forward(p, X, xm.astype(np.float64), xs.astype(np.float64))
    e32 = np.max(np.abs(out - ref32 * ys)) / np.max(np.abs(ref))
    print(f"inference: max diff {e:.2e} of the largest output (FP32-activation network: {e32:.2e})", flush=True)
    assert e < 1e-3
    assert e < 0.5 * e32
    # ragged row counts: a partial last tile, fewer rows than one tile
    for r in (1, 63, 65, 130):
        o2 = m.forward(Xd, r).download((r, MO.NO))
        assert np.max(np.abs(o2 - out[:r])) <= 1e-6 * np.max(np.abs(ref))
    g, loss = m.gradient(Xd, Yd, rows)
    lref, gref = MO.loss_and_grad(p, X, Y, xm.astype(np.float64), xs.astype(np.float64), ys.astype(np.float64))
    assert abs(loss - lref) < 1e-5 * lref
    o = 0
    for name, shp in MO.SHAPES:
        n = int(np.prod(shp))
        a, b = g[o:o + n], gref[o:o + n]
        e = np.max(np.abs(a - b)) / np.max(np.abs(b))
        print(f"{name}: max rel. diff {e:.2e}", flush=True)
        assert e < 1e-4, name
        o += n
    # two Adam steps of dg_mlp_train (the whole batch, no jitter) against
    # the oracle's Adam (oracle/mlp.py adam) fed with the GPU's gradients:
    # the parameters after each step within 1e-3 of the largest change (the
    # FP32 storage of the parameters rounds at ~1e-5 of the change)
    idx = hip.Buffer.from_array(np.arange(rows, dtype=np.int32), np.int32)
    lr = 1e-3
    pk, m1, m2 = p.astype(np.float64), 0.0, 0.0
    losses = []
    for k in (1, 2):
        gk, lk = m.gradient(Xd, Yd, rows)
        losses.append(lk)
        lt = m.train(Xd, Yd, rows, idx, rows, lr)
        assert abs(lt - lk) < 1e-6 * lk          # the step's own loss: the same batch
        pn, m1, m2 = MO.adam(pk, gk, m1, m2, k, lr)
        got = m.params().astype(np.float64)
        e = np.max(np.abs(got - pn)) / np.max(np.abs(pn - pk))
        print(f"Adam step {k}: max diff {e:.2e} of the largest parameter change", flush=True)
        assert e < 1e-3, k
        pk = got
    # and the steps move the loss down on this batch
    _, l2 = m.gradient(Xd, Yd, rows)
    assert l2 < losses[0]
    m.close()


def p1_faces_ref(U, ne):
    """face-mean jumps of a p = 1 Legendre state from the oracle's basis"""
    nx, ny, nz = ne
    x, w = O.gl_rule(3)
    phi = np.array([O.basis_1d("LegendreHex", 1, v)[0] for v in (-1.0, 1.0)])   # [side][function]
    mean = 0.5 * (w @ np.array([O.basis_1d("LegendreHex", 1, v)[0] for v in x]))  # (1/2) int phi_j
    Ub = U.reshape(nz, ny, nx, 2, 2, 2, 5)                           # k, j, i index order of b = i + 2j + 4k
    out = np.zeros((nz, ny, nx, 6, 5))

    def fm(u, d, side):
        # u [..., k, j, i, s]
        if d == 0:
            return np.einsum("...kjis,i,j,k->...s", u, phi[side], mean, mean)
        if d == 1:
            return np.einsum("...kjis,i,j,k->...s", u, mean, phi[side], mean)
        return np.einsum("...kjis,i,j,k->...s", u, mean, mean, phi[side])

    for d in range(3):
        for side in range(2):
            shift = 1 if side else -1
            axis = 2 - d                                               # array axes: z, y, x
            nb = np.roll(Ub, -shift, axis=axis)
            out[:, :, :, d * 2 + side] = fm(nb, d, 1 - side) - fm(Ub, d, side)
    return out.reshape(nx * ny * nz, 30)


def test_features_against_the_oracle(gpu):
    ne = (4, 3, 5)
    with dg_amd.Solver(dg_amd.case_toml(order=1, basis="LegendreHex", nelem=ne, timestep=0.01, ic=VORTEX[0],
                                        ic_data=VORTEX[1])) as g:
        g.init_state_from_fcn()
        U = g.get_state()
        X = closure.features(g, g.state_device_ptr()).download((U.shape[0], MO.NF))
        assert np.allclose(X[:, :40], U.reshape(-1, 40).astype(np.float32), rtol=0, atol=0)
        ref = p1_faces_ref(U, ne)
        assert np.max(np.abs(X[:, 40:] - ref)) < 1e-6 * max(np.max(np.abs(ref)), 1e-3)


def test_targets_against_the_oracle(gpu):
    ne = (3, 4, 3)
    fine_t = dg_amd.case_toml(order=3, basis="LegendreHex", flux="Roe", nelem=ne, timestep=0.01, ic=VORTEX[0],
                              ic_data=VORTEX[1])
    coarse_t = fine_t.replace("order = 3", "order = 1")
    with dg_amd.Solver(fine_t) as f, dg_amd.Solver(coarse_t) as c:
        f.init_state_from_fcn()
        Uf = f.get_state()
        E = Uf.shape[0]
        Uc = hip.Buffer(E * 40)
        Y = closure.target(f, f.state_device_ptr(), c, Uc).download((E, 40))
        of = O.Solver(O.make_params(order=3, basis="LegendreHex", flux="Roe", nelem=ne, lo=(-5, -5, -5),
                                    hi=(5, 5, 5), ic=VORTEX[0], ic_data=VORTEX[1]))
        oc = O.Solver(O.make_params(order=1, basis="LegendreHex", flux="Roe", nelem=ne, lo=(-5, -5, -5),
                                    hi=(5, 5, 5), ic=VORTEX[0], ic_data=VORTEX[1]))
        keep = [i + 4 * j + 16 * k for k in range(2) for j in range(2) for i in range(2)]
        trunc = Uf[:, keep, :]
        assert np.array_equal(Uc.download((E, 8, 5)), trunc)
        Df = of.inv_mass(1.0, of.residual(Uf))[:, keep, :]
        Dc = oc.inv_mass(1.0, oc.residual(trunc))
        ref = (Df - Dc).reshape(E, 40)
        assert np.max(np.abs(Y - ref)) < 1e-6 * np.max(np.abs(ref))


@pytest.mark.parametrize("flux", ["LaxFriedrichs", "HLLC", "Roe"])
def test_zero_network_steps_like_plain_rk4(gpu, flux):
    t = dg_amd.case_toml(order=1, basis="LegendreHex", flux=flux, nelem=(6, 5, 4), timestep=0.02, ic=VORTEX[0],
                         ic_data=VORTEX[1])
    m = closure.MLP(1)
    m.set_params(np.zeros(closure.num_params(), np.float32))
    with dg_amd.Solver(t) as a, dg_amd.Solver(t) as b:
        a.init_state_from_fcn()
        b.init_state_from_fcn()
        nfb = closure.step(a, m, 1e30, 5)
        b.take_time_step(0.0, 5)
        assert nfb == 0
        assert a.time == b.time
        Ua, Ub = a.get_state(), b.get_state()
        assert np.max(np.abs(Ua - Ub)) < 1e-13 * np.max(np.abs(Ub))
        # a bound of 0 turns every nonzero correction off: the same result
    m.close()


def closure_rk4_ref(o, U, ne, dt, net):
    """one RK4 step of dU/dt = M^-1 R(U) + C(U) on the oracle: C from the
    oracle features (the p = 1 coefficients and face-mean jumps) and the
    oracle network with bf16 layer inputs, in FP32 like the GPU's C"""
    p, xm, xs, ys = net
    E = U.shape[0]

    def f(V):
        X = np.concatenate([V.reshape(E, 40), p1_faces_ref(V, ne)], axis=1).astype(np.float32)
        C = (MO.forward(p, X, xm.astype(np.float64), xs.astype(np.float64), round_acts=True)[0] * ys)
        return o.inv_mass(1.0, o.residual(V)) + C.astype(np.float32).reshape(E, 8, 5)

    k1 = f(U)
    k2 = f(U + 0.5 * dt * k1)
    k3 = f(U + 0.5 * dt * k2)
    k4 = f(U + dt * k3)
    return U + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)


def test_closure_step_against_the_oracle(gpu):
    """A random network (outputs of order 1, so the correction moves the
    state by ~1e-2 in one step) through one corrected RK4 step against the
    oracle's: bar 1e-4 of the state (the correction is FP32 with bf16 layer
    inputs); with bound 0 every row falls back at every stage and the step
    is plain RK4's."""
    ne = (4, 3, 5)
    dt = 0.02
    t = dg_amd.case_toml(order=1, basis="LegendreHex", flux="HLLC", nelem=ne, timestep=dt, ic=VORTEX[0],
                         ic_data=VORTEX[1])
    p, xm, xs, ys = rand_net(2)
    p = p * 0.5
    net = (p, xm, xs, ys)
    m = closure.MLP(0)
    m.set_params(p)
    m.set_scaling(xm, xs, ys)
    o = O.Solver(O.make_params(order=1, basis="LegendreHex", flux="HLLC", nelem=ne, lo=(-5, -5, -5), hi=(5, 5, 5),
                               ic=VORTEX[0], ic_data=VORTEX[1]))
    with dg_amd.Solver(t) as a, dg_amd.Solver(t) as b:
        a.init_state_from_fcn()
        b.init_state_from_fcn()
        U0 = a.get_state()
        E = U0.shape[0]
        nfb = closure.step(a, m, 1e30, 1)
        assert nfb == 0
        ref = closure_rk4_ref(o, U0, ne, dt, net)
        b.take_time_step(0.0, 1)
        plain = b.get_state()
        U = a.get_state()
        e = np.max(np.abs(U - ref)) / np.max(np.abs(ref))
        moved = np.max(np.abs(ref - plain)) / np.max(np.abs(ref))
        print(f"corrected step: rel. diff vs oracle {e:.2e}; the correction moved the state by {moved:.2e}",
              flush=True)
        assert moved > 1e-3
        assert e < 1e-4
        # bound 0: all rows fall back (4 stages x E rows), plain RK4 results
        a.set_state(U0)
        a.set_time(0.0)
        nfb = closure.step(a, m, 0.0, 1)
        assert nfb == 4 * E
        assert np.max(np.abs(a.get_state() - plain)) < 1e-13 * np.max(np.abs(plain))
    m.close()


def test_trained_network_gain(gpu):
    """What the trained network gains: trained along coarse runs beside
    fine ones (two rollout rounds), on held-out seeds it removes most of the
    coarse tendency's error a priori, and a posteriori the corrected p = 1
    run to t = 0.5 stays finite without any fallback and ends closer to the
    fine run than the plain p = 1 run (16^3 box, p = 3 fine runs; measured
    0.66x the plain density error at this size)."""
    n = 16
    m, (h1, h2), rows = closure.train_rollouts(list(range(1000, 1008)), n, fine_order=3, steps=24,
                                                tau_steps=4.0, epochs=6, bound=2.0)
    print("losses", h1, h2, flush=True)
    assert h1[-1] < h1[0] and h2[-1] < h2[0]
    Xv, Yv, rv = closure.generate_rollout([7, 8], n, fine_order=3, steps=12, tau_steps=4.0)
    mse_net, mse_zero = closure.a_priori(m, Xv, Yv, rv)
    print(f"held-out MSE: network {mse_net:.3e}, zero correction {mse_zero:.3e}", flush=True)
    assert mse_net < 0.5 * mse_zero
    ev = closure.evaluate(m, [1, 2, 3, 4], n, 0.5, fine_order=3, bounds=(None,), compare_orders=())
    p1 = np.mean([r["p1_rho_error"] for r in ev])
    cl = np.mean([r["closure_rho_error"] for r in ev])
    print(f"density error: plain p=1 {p1:.4e}, closure {cl:.4e} ({cl / p1:.3f}x)", flush=True)
    assert np.isfinite(cl) and sum(r["closure_fallback_elements"] for r in ev) == 0
    assert cl < 0.85 * p1
    m.close()


def test_rollout_element_subsampling(gpu):
    """generate_rollout with stride s keeps, at coarse step k of seed sd,
    the rows of elements j, j + s, ... with j = (5 k + sd) mod s of the full
    rollout: the same runs, bit-identical rows."""
    n, steps, seeds = 8, 3, [11, 12]
    X1, Y1, r1 = closure.generate_rollout(seeds, n, fine_order=2, steps=steps)
    X4, Y4, r4 = closure.generate_rollout(seeds, n, fine_order=2, steps=steps, stride=4)
    E = n ** 3
    assert r1 == len(seeds) * steps * E and r4 == r1 // 4
    Xf, Yf = X1.download((r1, MO.NF)), Y1.download((r1, MO.NO))
    Xs, Ys = X4.download((r4, MO.NF)), Y4.download((r4, MO.NO))
    for i, sd in enumerate(seeds):
        for k in range(steps):
            j = (5 * k + sd) % 4
            full = slice((i * steps + k) * E, (i * steps + k + 1) * E)
            sub = slice((i * steps + k) * (E // 4), (i * steps + k + 1) * (E // 4))
            assert np.array_equal(Xs[sub], Xf[full][j::4])
            assert np.array_equal(Ys[sub], Yf[full][j::4])
    for b in (X1, Y1, X4, Y4):
        b.free()

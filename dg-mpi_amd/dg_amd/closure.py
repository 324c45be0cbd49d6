"""Learned closure for coarse (p = 1 Legendre) periodic-box runs over the C
ABI (dg_mlp_*, dg_closure_*): a per-element MLP, 70 -> 128 -> 128 -> 40,
GELU, BF16 weights with FP32 accumulation, trained by the library's own
forward / backward / Adam kernels.

    m = MLP(seed)
    m.set_scaling(x_mean, x_scale, y_scale)
    loss = m.train(X, Y, idx, batch, lr)          # device buffers (hip.Buffer)
    X = features(coarse, U_ptr)                    # [E][70] float32 buffer
    Y = target(fine, Uf_ptr, coarse, Uc)           # [E][40] float32 buffer
    nfb = step(coarse, m, bound, nsteps)           # corrected RK4 steps
"""
import ctypes as C

import numpy as np

from . import DGError, lib
from . import hip

NF, NH, NO = 70, 128, 40


def _check(rc, m=None):
    if rc != 0:
        raise DGError(rc, lib().dg_mlp_last_error(m._h if m is not None else None).decode())


def num_params():
    n = C.c_long()
    _check(lib().dg_mlp_num_params(C.byref(n)))
    return n.value


class MLP:
    def __init__(self, seed=0):
        h = C.c_void_p()
        _check(lib().dg_mlp_create(seed, C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().dg_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.empty(num_params(), dtype=np.float32)
        _check(lib().dg_mlp_get_params(self._h, p.ctypes.data), self)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == num_params()
        _check(lib().dg_mlp_set_params(self._h, p.ctypes.data), self)

    def set_scaling(self, x_mean, x_scale, y_scale):
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (x_mean, x_scale, y_scale)]
        assert a[0].size == NF and a[1].size == NF and a[2].size == NO
        _check(lib().dg_mlp_set_scaling(self._h, a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data), self)
        self.scaling = a

    def forward(self, X, rows):
        Y = hip.Buffer(rows * NO, np.float32)
        _check(lib().dg_mlp_forward(self._h, X.addr, Y.addr, rows), self)
        return Y

    def gradient(self, X, Y, rows):
        g = np.empty(num_params(), dtype=np.float32)
        loss = C.c_double()
        _check(lib().dg_mlp_gradient(self._h, X.addr, Y.addr, rows, g.ctypes.data, C.byref(loss)), self)
        return g, loss.value

    def train(self, X, Y, nrows, idx, batch, lr, noise=0.0):
        loss = C.c_double()
        _check(lib().dg_mlp_train(self._h, X.addr, Y.addr, nrows, idx.addr, idx.n, batch, C.c_float(lr),
                                  C.c_float(noise), C.byref(loss)), self)
        return loss.value


def features(solver, U_ptr, out=None):
    E = solver.info.num_elems
    X = out if out is not None else hip.Buffer(E * NF, np.float32)
    _check(lib().dg_closure_features(solver._h, C.c_void_p(U_ptr), X.addr))
    return X


def target(fine, Uf_ptr, coarse, Uc, Y=None):
    E = coarse.info.num_elems
    Y = Y if Y is not None else hip.Buffer(E * NO, np.float32)
    _check(lib().dg_closure_target(fine._h, C.c_void_p(Uf_ptr), coarse._h, Uc.addr, Y.addr))
    return Y


def target_at(fine, Uf_ptr, coarse, Uc_ptr, tau, Y):
    _check(lib().dg_closure_target_at(fine._h, C.c_void_p(Uf_ptr), coarse._h, C.c_void_p(Uc_ptr), tau, Y.addr))
    return Y


def rollout_rows(nseeds, n, steps, stride=1):
    """rows that generate_rollout writes: E / stride per coarse step"""
    return nseeds * steps * (n ** 3 // stride)


def generate_rollout(seeds, n, fine_order=4, steps=24, tau_steps=4.0, model=None, bound=None, out=None, log=None,
                     stride=1):
    """Training rows along coarse runs beside fine ones: per seed the fine
    run and a coarse p = 1 run (plain, or with `model` and its fallback
    bound) advance together, the coarse step an integer number of fine
    steps; after each coarse step, the features of the coarse state and the
    target of dg_closure_target_at with tau = tau_steps coarse steps.
    stride > 1 (a divisor of E) keeps every stride-th element of each step,
    starting at an element that rotates with the step and the seed, so that
    many seeds fit in memory with their whole time history.
    out = (X, Y, first_row): write into existing buffers from that row."""
    from . import Solver
    E = n ** 3
    assert stride >= 1 and E % stride == 0
    per_step = E // stride
    rows = rollout_rows(len(seeds), n, steps, stride)
    if out is None:
        X = hip.Buffer(rows * NF, np.float32)
        Y = hip.Buffer(rows * NO, np.float32)
        r = 0
    else:
        X, Y, r = out
    Xs = hip.Buffer(E * NF, np.float32) if stride > 1 else None
    Ys = hip.Buffer(E * NO, np.float32) if stride > 1 else None
    for sd in seeds:
        data = mix_ic(np.random.default_rng(sd))
        with Solver(case(fine_order, n, data, CFL[fine_order])) as f, Solver(case(1, n, data, CFL[1])) as c:
            f.init_state_from_fcn()
            c.init_state_from_fcn()
            dtc = c.timestep
            mult = int(np.ceil(dtc / f.timestep))
            dtf = dtc / mult
            for k in range(steps):
                f.take_time_step(dtf, mult)
                if model is None:
                    c.take_time_step(dtc, 1)
                else:
                    step(c, model, 1e30 if bound is None else bound * model.target_rms, 1, dtc)
                if stride == 1:
                    target_at(f, f.state_device_ptr(), c, c.state_device_ptr(), tau_steps * dtc,
                              _At(Y.addr + r * NO * 4))
                    features(c, c.state_device_ptr(), _At(X.addr + r * NF * 4))
                else:
                    target_at(f, f.state_device_ptr(), c, c.state_device_ptr(), tau_steps * dtc, Ys)
                    features(c, c.state_device_ptr(), Xs)
                    j = (k * 5 + sd) % stride
                    X.copy_rows_from(r, Xs, j, NF, per_step, stride)
                    Y.copy_rows_from(r, Ys, j, NO, per_step, stride)
                r += per_step
        if log:
            log(f"rollout rows: seed {sd} done ({r} rows)")
    hip.synchronize()
    if Xs is not None:
        Xs.free()
        Ys.free()
    return X, Y, rows


def train_rollouts(seeds, n, fine_order=4, steps=24, tau_steps=4.0, epochs=6, bound=2.0, log=None, stride=1,
                   rounds=2, noise=0.0):
    """Rounds of rollout training (DAgger, Ross et al. 2011): rows along
    plain coarse runs and a first fit; then, per further round, rows along
    coarse runs corrected by the latest network (fallback bound in
    target-RMS units) and a fit on all rows so far (noise: Gaussian feature
    jitter of the fits, normalised units; stride: element subsampling of
    generate_rollout).  Returns the last network, the loss histories and
    the row count."""
    per = rollout_rows(len(seeds), n, steps, stride)
    X = hip.Buffer(rounds * per * NF, np.float32)
    Y = hip.Buffer(rounds * per * NO, np.float32)
    model, hist = None, []
    for rd in range(rounds):
        generate_rollout(seeds, n, fine_order, steps, tau_steps, model=model, bound=bound, out=(X, Y, rd * per),
                         log=log, stride=stride)
        m, h = fit(X, Y, (rd + 1) * per, epochs=epochs, batch=4096, lr=2e-3, seed=rd, noise=noise, log=log)
        if model is not None:
            model.close()
        model = m
        hist.append(h)
    X.free()
    Y.free()
    return model, tuple(hist), rounds * per


def step(solver, mlp, bound, nsteps, dt=0.0):
    n = C.c_long()
    _check(lib().dg_closure_step(solver._h, mlp._h, dt, bound, nsteps, C.byref(n)))
    solver.info = solver._info()
    return n.value


def step_noalloc(solver, mlp, bound, nsteps, dt=0.0):
    """step() through dg_closure_step_noalloc (the library's static device
    scratch, at most 32^3 elements; no device allocation)."""
    n = C.c_long()
    _check(lib().dg_closure_step_noalloc(solver._h, mlp._h, dt, bound, nsteps, C.byref(n)))
    solver.info = solver._info()
    return n.value


# ---------------------------------------------------------------------------
# Training data, fitting and evaluation (the workflow of the closure)
# ---------------------------------------------------------------------------

class _At:
    """A device address inside a larger buffer (row offset of X or Y)."""

    def __init__(self, addr):
        self.addr = addr


def mix_ic(rng, L=10.0):
    """InitialCondition.data of a VortexPulseMix: 1-4 vortices at random
    centres with strengths in [1, 5], a Gaussian pressure pulse of amplitude
    0.01-0.05 and width 1, free stream (ub, vb) in [-0.5, 0.5]^2."""
    nv = int(rng.integers(1, 5))
    d = [float(rng.uniform(-0.5, 0.5)), float(rng.uniform(-0.5, 0.5))]
    for k in range(4):
        c = rng.uniform(-L / 2, L / 2, 2)
        d += [float(c[0]), float(c[1]), float(rng.uniform(1.0, 5.0)) if k < nv else 0.0]
    pc = rng.uniform(-L / 2, L / 2, 3)
    d += [float(rng.uniform(0.01, 0.05)), float(pc[0]), float(pc[1]), float(pc[2]), 1.0]
    return d


def case(order, n, data, cfl):
    from . import case_toml
    return case_toml(order=order, basis="LegendreHex", flux="Roe", nelem=(n, n, n), cfl=cfl, end=1.0,
                     ic="VortexPulseMix", ic_data=data)


CFL = {1: 0.1, 2: 0.1, 3: 0.06, 4: 0.04}


def generate(seeds, n, fine_order=4, snapshots=10, every=4):
    """Training rows from FP64 fine runs: for each seed, `snapshots` fine
    states `every` fine steps apart, each giving E rows of features of the
    projected coarse state and targets.  Returns device buffers X, Y and the
    row count."""
    from . import Solver
    E = n ** 3
    rows = len(seeds) * snapshots * E
    X = hip.Buffer(rows * NF, np.float32)
    Y = hip.Buffer(rows * NO, np.float32)
    Uc = hip.Buffer(E * 40)
    r = 0
    for sd in seeds:
        data = mix_ic(np.random.default_rng(sd))
        with Solver(case(fine_order, n, data, CFL[fine_order])) as f, Solver(case(1, n, data, CFL[1])) as c:
            f.init_state_from_fcn()
            for k in range(snapshots):
                target(f, f.state_device_ptr(), c, Uc, _At(Y.addr + r * NO * 4))
                features(c, Uc.addr, _At(X.addr + r * NF * 4))
                r += E
                f.take_time_step(0.0, every)
    hip.synchronize()
    return X, Y, rows


def scaling_from(X, Y, rows, sample=200000, seed=0):
    """feature mean and 1/std, target std, and the RMS of the targets'
    per-element 2-norm (the unit of the fallback bound)"""
    rng = np.random.default_rng(seed)
    # 16 blocks of rows spread over the first `rows` rows (the buffers may hold more)
    blk = max(1, min(rows, sample) // 16)
    starts = np.linspace(0, rows - blk, 16).astype(np.int64)
    Xh = np.concatenate([X.download_range(s0 * NF, blk * NF) for s0 in starts]).reshape(-1, NF)
    Yh = np.concatenate([Y.download_range(s0 * NO, blk * NO) for s0 in starts]).reshape(-1, NO)
    pick = np.arange(len(Xh))
    xm = Xh[pick].mean(0)
    xs = 1.0 / np.maximum(Xh[pick].std(0), 1e-6)
    ys = np.maximum(Yh[pick].std(0), 1e-8)
    rms = float(np.sqrt(np.mean(np.sum(Yh[pick].astype(np.float64) ** 2, axis=1))))
    return xm, xs, ys, rms


def fit(X, Y, rows, epochs=10, batch=4096, lr=1e-3, seed=0, noise=0.0, log=None):
    m = MLP(seed)
    xm, xs, ys, rms = scaling_from(X, Y, rows, seed=seed)
    m.set_scaling(xm, xs, ys)
    m.target_rms = rms
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(rows).astype(np.int32)
        idx = hip.Buffer.from_array(perm, np.int32)
        rate = lr * 0.5 * (1 + np.cos(np.pi * ep / epochs))      # cosine decay
        loss = m.train(X, Y, rows, idx, batch, rate, noise)
        idx.free()
        hist.append(loss)
        if log:
            log(f"epoch {ep}: loss {loss:.4e} (lr {rate:.2e})")
    return m, hist


def l2_p1(U, n, L=10.0):
    """Volume-averaged L2 norm per variable of p = 1 Legendre coefficients
    [E][8][5] on the periodic box of n^3 elements of edge L (unnormalised
    Legendre polynomials: int P0^2 = 2, int P1^2 = 2/3 on [-1, 1])."""
    w = np.array([np.prod([2.0 / (2 * ((b >> d) & 1) + 1) for d in range(3)]) for b in range(8)])
    jac = (L / n / 2.0) ** 3
    return np.sqrt(np.einsum("ebs,b->s", np.asarray(U).reshape(-1, 8, 5) ** 2, w) * jac / L ** 3)


def a_priori(m, X, Y, rows):
    """normalised mean squared error of the network on rows (X, Y) and of
    the zero correction (the plain coarse tendency): what the network
    removes of the coarse tendency's error"""
    _, loss = m.gradient(X, Y, rows)
    Yh = Y.download((rows, NO), rows * NO).astype(np.float64) / m.scaling[2].astype(np.float64)
    return loss, float(np.mean(Yh * Yh))


def evaluate(m, seeds, n, t_end, fine_order=4, bounds=(None,), relax=(1.0,), compare_orders=(3,)):
    """Held-out runs to t_end: the fine reference, plain p = 1, p = 1 with
    the closure for each fallback bound (in units of the training targets'
    RMS element norm; None: no fallback), and plain runs of compare_orders;
    errors of the p = 1 projections against the fine one's (density,
    volume-averaged L2) and wall seconds per simulated time unit."""
    import time
    from . import Solver
    keep = {}
    out = []
    for sd in seeds:
        data = mix_ic(np.random.default_rng(sd))
        res = {"seed": int(sd)}

        def run(order, fn):
            s = Solver(case(order, n, data, CFL[order]))
            s.init_state_from_fcn()
            nst = max(1, int(round(t_end / s.timestep)))
            dt = t_end / nst
            hip.synchronize()
            t0 = time.perf_counter()
            extra = fn(s, dt, nst)
            hip.synchronize()
            wall = time.perf_counter() - t0
            U = s.get_state()
            nb = (order + 1) ** 3
            k = [i + (order + 1) * j + (order + 1) ** 2 * kk for kk in range(2) for j in range(2) for i in range(2)]
            s.close()
            return U.reshape(-1, nb, 5)[:, k, :], wall / t_end, extra

        def plain(s, dt, nst):
            s.take_time_step(dt, nst)

        ref, res["fine_wall_per_time"], _ = run(fine_order, plain)
        p1, res["p1_wall_per_time"], _ = run(1, plain)
        res["p1_rho_error"] = float(l2_p1(p1 - ref, n)[0])
        xm, xs, ys = m.scaling
        for bnd in bounds:
            for a in relax:
                # the correction scaled by a (the output scale times a)
                m.set_scaling(xm, xs, ys * a)
                b = 1e30 if bnd is None else bnd * m.target_rms
                cl, wall, nfb = run(1, lambda s, dt, nst: step(s, m, b, nst, dt))
                m.set_scaling(xm, xs, ys)
                key = ("closure" if bnd is None else f"closure_b{bnd:g}") + ("" if a == 1.0 else f"_r{a:g}")
                res[f"{key}_wall_per_time"] = wall
                res[f"{key}_fallback_elements"] = nfb
                res[f"{key}_rho_error"] = float(l2_p1(cl - ref, n)[0])
        for q in compare_orders:
            pq, res[f"p{q}_wall_per_time"], _ = run(q, plain)
            res[f"p{q}_rho_error"] = float(l2_p1(pq - ref, n)[0])
        out.append(res)
    return out

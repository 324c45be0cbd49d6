"""Pins of the closure network's CPU restatement (oracle/mlp.py): the
backward pass against central finite differences of the loss (weights not
rounded, so the loss is smooth), the bfloat16 rounding against values it
must produce, the GELU derivative against differences of GELU, GELU itself
against x Phi(x) at known values of the normal CDF and against torch's exact
GELU, and the Adam update against its closed forms (first step, constant
gradient) and torch.optim.Adam on a varying gradient sequence."""
import numpy as np
import pytest

import mlp as M


def data(rows=12, seed=0):
    rng = np.random.default_rng(seed)
    p = np.concatenate([rng.normal(0, 0.3, int(np.prod(s))) for _, s in M.SHAPES]).astype(np.float32)
    X = rng.normal(0, 1, (rows, M.NF)).astype(np.float32)
    Y = rng.normal(0, 1, (rows, M.NO)).astype(np.float32)
    xm = rng.normal(0, 0.1, M.NF)
    xs = rng.uniform(0.5, 2.0, M.NF)
    ys = rng.uniform(0.5, 2.0, M.NO)
    return p, X, Y, xm, xs, ys


def test_backward_against_finite_differences():
    p, X, Y, xm, xs, ys = data()
    p = p.astype(np.float64)
    _, g = M.loss_and_grad(p, X, Y, xm, xs, ys, round_weights=False)
    rng = np.random.default_rng(1)
    # a sample of parameters from every block
    o = 0
    picks = []
    for _, shp in M.SHAPES:
        n = int(np.prod(shp))
        picks += list(o + rng.choice(n, 6, replace=False))
        o += n
    h = 1e-5
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, X, Y, xm, xs, ys, False)[0] - M.loss_and_grad(pm, X, Y, xm, xs, ys, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.1415926, 1e-20], dtype=np.float32)
    r = M.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                       # tie, to the even neighbour
    assert r[2] == 1.0 + 2 ** -7             # above the tie
    assert r[3] == -2.5
    assert r[4] == np.float32(3.140625)
    assert abs(r[5] - 1e-20) < 1e-22


def test_gelu_derivative():
    z = np.linspace(-5, 5, 101)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_d(z))) < 1e-8


def test_activation_rounding_stays_within_its_propagated_bound():
    """forward(round_acts=True) rounds the input of each layer to bfloat16
    (8 significant bits: relative error at most 2^-8 per value).
    First-order propagation through |W| and max |GELU'| < 1.13 bounds the
    output change; the change must stay under that bound; its RMS must
    match the random-rounding estimate (independent errors, uniform within
    half a bf16 ulp: variance ulp(a)^2 / 12, propagated through W^2 and
    GELU'(z)^2) within 30 %, which fails if a layer's rounding is skipped or
    doubled; and with activations that are already bfloat16 numbers (bf16
    inputs, identity scaling, the hidden weights zero) nothing changes."""
    p, X, _, xm, xs, _ = data(rows=64, seed=5)
    O, (A0, Z1, H1, Z2, H2) = M.forward(p, X, xm, xs)
    Or, _ = M.forward(p, X, xm, xs, round_acts=True)
    P = M.unpack(p)
    u, g = 2.0 ** -8, 1.13
    W1, W2, W3 = (np.abs(M.bf16(P[k])) for k in ("W1", "W2", "W3"))
    e1 = (u * np.abs(A0)) @ W1
    e2 = (g * e1 + u * np.abs(H1)) @ W2
    e3 = (g * e2 + u * np.abs(H2)) @ W3
    d = np.abs(Or - O)
    assert np.all(d <= 1.05 * e3 + 1e-12)

    def var(a):                                   # a = m 2^e, m in [0.5, 1): ulp = 2^(e - 8)
        _, e = np.frexp(np.asarray(a, np.float32))
        return np.where(a != 0, np.ldexp(1.0, e - 8) ** 2 / 12.0, 0.0)

    v1 = var(A0) @ W1 ** 2
    v2 = (M.gelu_d(Z1) ** 2 * v1 + var(H1)) @ W2 ** 2
    v3 = (M.gelu_d(Z2) ** 2 * v2 + var(H2)) @ W3 ** 2
    ratio = np.sqrt(np.mean(d ** 2) / np.mean(v3))
    print(f"activation rounding: measured / estimated RMS {ratio:.3f}")
    assert 0.7 < ratio < 1.3
    # bf16-exact features, identity scaling, zero hidden weights: the only
    # rounded values are already representable, so the outputs agree exactly
    q = p.copy()
    o = 0
    for name, shp in M.SHAPES:
        n = int(np.prod(shp))
        if name in ("W2", "W3"):
            q[o:o + n] = 0.0
        o += n
    Xb = M.bf16(X).astype(np.float32)
    a, _ = M.forward(q, Xb, np.zeros(M.NF), np.ones(M.NF))
    b, _ = M.forward(q, Xb, np.zeros(M.NF), np.ones(M.NF), round_acts=True)
    assert np.array_equal(a, b)


def test_gelu_known_values_and_torch():
    # GELU(x) = x Phi(x): Phi(1) = 0.8413447460685429, Phi(2) = 0.9772498680518208,
    # Phi(-0.5) = 0.3085375387259869 (normal CDF tables)
    assert abs(M.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15
    assert abs(M.gelu(np.array([2.0]))[0] - 2.0 * 0.9772498680518208) < 1e-15
    assert abs(M.gelu(np.array([-0.5]))[0] + 0.5 * 0.3085375387259869) < 1e-15
    assert M.gelu(np.array([0.0]))[0] == 0.0
    torch = pytest.importorskip("torch")
    z = np.linspace(-6, 6, 241)
    ref = torch.nn.functional.gelu(torch.tensor(z, dtype=torch.float64), approximate="none").numpy()
    assert np.max(np.abs(M.gelu(z) - ref)) < 1e-15
    # the tanh approximation differs from the exact form by ~1e-4: a wrong form is caught
    tanh = torch.nn.functional.gelu(torch.tensor(z, dtype=torch.float64), approximate="tanh").numpy()
    assert np.max(np.abs(M.gelu(z) - tanh)) > 1e-5


def test_adam_closed_forms():
    rng = np.random.default_rng(5)
    p0 = rng.normal(0, 1, 50)
    g = rng.normal(0, 1, 50)
    lr, eps = 1e-3, 1e-8
    # first step from zero moments: the bias-corrected moments are g and g^2
    p1, m1, m2 = M.adam(p0, g, 0.0, 0.0, 1, lr)
    assert np.allclose(p1, p0 - lr * g / (np.abs(g) + eps), rtol=0, atol=1e-15)
    assert np.allclose(m1, 0.1 * g) and np.allclose(m2, 0.001 * g * g)
    # a constant gradient keeps m_hat = g, v_hat = g^2 at every step
    p, m1, m2 = p0, 0.0, 0.0
    for k in range(1, 8):
        p, m1, m2 = M.adam(p, g, m1, m2, k, lr)
    assert np.allclose(p, p0 - 7 * lr * g / (np.abs(g) + eps), rtol=0, atol=1e-13)


def test_adam_against_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(6)
    p0 = rng.normal(0, 1, 40)
    grads = [rng.normal(0, 1, 40) * s for s in (1.0, -3.0, 0.2, 5.0, 1.0)]
    t = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([t], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m1, m2 = p0, 0.0, 0.0
    for k, g in enumerate(grads, 1):
        opt.zero_grad()
        t.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
        p, m1, m2 = M.adam(p, g, m1, m2, k, 1e-2)
        assert np.max(np.abs(p - t.detach().numpy())) < 1e-13, k

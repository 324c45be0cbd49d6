"""Pins of the closure network's CPU restatement (oracle/mlp.py): the
backward pass against central finite differences of the loss (weights not
rounded, so the loss is smooth), the bfloat16 rounding against values it
must produce, and the GELU derivative against differences of GELU."""
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

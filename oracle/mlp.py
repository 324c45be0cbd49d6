"""CPU restatement of the learned-closure network (test infrastructure only:
the checker of the GPU kernels, never the product path).

70 -> 128 -> 128 -> 40, GELU (exact, erf form), weights rounded to bfloat16
(round to nearest even) for the forward pass and for the backward pass to
the inputs, FP32 master weights for the weight gradients; inputs normalised
as (x - mean) * scale, outputs in units of the target scale; loss = mean
squared error over rows and outputs.  round_acts=True also rounds each
layer's input (normalised features, hidden activations) to bfloat16, as the
inference kernel (dg_mlp_forward and the closure step) feeds the matrix
cores; training keeps FP32 activations.  Computed in float64 here, so the
GPU's FP32 accumulation differs by rounding only.  Pinned by finite
differences of the loss and by the size of the activation rounding
(tests/test_oracle_mlp.py); GELU against known values of x Phi(x) and
torch's exact GELU; the Adam update (adam) against its closed forms and
torch.optim.Adam."""
import numpy as np
from scipy.special import erf

NF, NH, NO = 70, 128, 40
SHAPES = [("W1", (NF, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NO)), ("b3", (NO,))]


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    assert o == len(p)
    return out


def bf16(a):
    """float32 values rounded to bfloat16 (nearest even), as float64"""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(p, X, x_mean, x_scale, round_weights=True, round_acts=False):
    P = unpack(p)
    r = bf16 if round_weights else (lambda w: np.asarray(w, dtype=np.float64))
    ra = bf16 if round_acts else (lambda a: a)
    A0 = (np.asarray(X, np.float64) - x_mean) * x_scale
    Z1 = ra(A0) @ r(P["W1"]) + P["b1"]
    H1 = gelu(Z1)
    Z2 = ra(H1) @ r(P["W2"]) + P["b2"]
    H2 = gelu(Z2)
    O = ra(H2) @ r(P["W3"]) + P["b3"]
    return O, (A0, Z1, H1, Z2, H2)


def loss_and_grad(p, X, Y, x_mean, x_scale, y_scale, round_weights=True):
    P = unpack(p)
    r = bf16 if round_weights else (lambda w: np.asarray(w, dtype=np.float64))
    O, (A0, Z1, H1, Z2, H2) = forward(p, X, x_mean, x_scale, round_weights)
    M = len(X)
    D = O - np.asarray(Y, np.float64) / y_scale
    loss = float(np.mean(D * D))
    G3 = 2.0 * D / (M * NO)
    g = {"W3": H2.T @ G3, "b3": G3.sum(0)}
    G2 = (G3 @ r(P["W3"]).T) * gelu_d(Z2)
    g["W2"], g["b2"] = H1.T @ G2, G2.sum(0)
    G1 = (G2 @ r(P["W2"]).T) * gelu_d(Z1)
    g["W1"], g["b1"] = A0.T @ G1, G1.sum(0)
    return loss, np.concatenate([g[name].ravel() for name, _ in SHAPES])


def adam(p, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) of parameters p with
    gradient g and moments m1, m2 after `step` updates (step >= 1 is this
    update's count), in float64.  Returns (p, m1, m2)."""
    g = np.asarray(g, np.float64)
    m1 = b1 * np.asarray(m1, np.float64) + (1.0 - b1) * g
    m2 = b2 * np.asarray(m2, np.float64) + (1.0 - b2) * g * g
    mhat = m1 / (1.0 - b1 ** step)
    vhat = m2 / (1.0 - b2 ** step)
    return np.asarray(p, np.float64) - lr * mhat / (np.sqrt(vhat) + eps), m1, m2

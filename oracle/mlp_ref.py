"""float64 numpy reference of the surrogate MLP (16 -> 64 -> 64 -> 4, GELU)
used to check the HIP forward/backward kernels.  TEST INFRASTRUCTURE ONLY.

Parameter packing (the C ABI's): W1[64][16] b1[64] W2[64][64] b2[64]
W3[4][64] b3[4].  Loss: mean of squared errors over batch x 4 outputs.
"""
import math

import numpy as np

IN, H, OUT = 16, 64, 4
SHAPES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
          ("W3", (OUT, H)), ("b3", (OUT,))]
NUM_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)

_erf = np.vectorize(math.erf)


def unpack(params):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(params[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[n], dtype=np.float64).ravel() for n, _ in SHAPES])


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    cdf = 0.5 * (1.0 + _erf(z / math.sqrt(2.0)))
    return cdf + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(params, x):
    p = unpack(params)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ p["W1"].T + p["b1"]
    h1 = gelu(z1)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = gelu(z2)
    y = h2 @ p["W3"].T + p["b3"]
    return y, (x, z1, h1, z2, h2)


def loss_and_grad(params, x, t):
    p = unpack(params)
    y, (x, z1, h1, z2, h2) = forward(params, x)
    B = x.shape[0]
    e = y - np.asarray(t, dtype=np.float64)
    loss = float((e * e).mean())
    d3 = 2.0 * e / (B * OUT)
    d2 = (d3 @ p["W3"]) * gelu_grad(z2)
    d1 = (d2 @ p["W2"]) * gelu_grad(z1)
    g = {"W1": d1.T @ x, "b1": d1.sum(0), "W2": d2.T @ h1, "b2": d2.sum(0),
         "W3": d3.T @ h2, "b3": d3.sum(0)}
    return loss, pack(g)


def init_params(seed=0):
    """He-style initialisation (fan-in), zero biases."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            d[name] = rng.standard_normal(shape) * math.sqrt(2.0 / shape[1])
        else:
            d[name] = np.zeros(shape)
    return pack(d).astype(np.float32)

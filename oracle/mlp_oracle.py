"""FP64 NumPy restatement of the surrogate MLP (16-128-128-128-7, SiLU).

TEST INFRASTRUCTURE ONLY: the checker for the HIP forward and gradient
kernels (materialpointsolver.jl_amd/csrc/mpm_mlp.hpp).  Pinned by central
finite differences in tests/test_oracle_mlp.py.

Parameter layout (flat FP32 vector, shared with the library): W1[128][16] b1
W2[128][128] b2 W3[128][128] b3 W4[7][128] b4, then x_mean[16] x_istd[16]
y_scale[7].  The network maps standardised inputs (x - x_mean) * x_istd to
normalised targets y / y_scale; the loss is the mean squared error over all
B * 7 normalised outputs.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 16, 128, 7
SHAPES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
          ("W3", (NH, NH)), ("b3", (NH,)), ("W4", (NOUT, NH)), ("b4", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
NTOTAL = NPARAM + NIN + NIN + NOUT


def unpack(p):
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    out["x_mean"] = np.asarray(p[o:o + NIN], dtype=np.float64)
    out["x_istd"] = np.asarray(p[o + NIN:o + 2 * NIN], dtype=np.float64)
    out["y_scale"] = np.asarray(p[o + 2 * NIN:o + 2 * NIN + NOUT], dtype=np.float64)
    return out


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, rounding=False):
    """Outputs in target units (y = network output * y_scale), shape (B, 7).
    rounding=True rounds the operands the matrix cores see to bfloat16."""
    q = unpack(p)
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    h = r((np.asarray(X, dtype=np.float64) - q["x_mean"]) * q["x_istd"])
    for k in (1, 2, 3):
        h = r(silu(h @ r(q[f"W{k}"]).T + q[f"b{k}"]))
    o = h @ r(q["W4"]).T + q["b4"]
    return o * q["y_scale"]


def loss_and_grad(p, X, Y):
    """Mean squared error over normalised targets and its gradient with respect
    to the NPARAM trainable parameters (FP64, no rounding)."""
    q = unpack(p)
    x = (np.asarray(X, dtype=np.float64) - q["x_mean"]) * q["x_istd"]
    yn = np.asarray(Y, dtype=np.float64) / q["y_scale"]
    zs, hs = [], [x]
    h = x
    for k in (1, 2, 3):
        z = h @ q[f"W{k}"].T + q[f"b{k}"]
        zs.append(z)
        h = silu(z)
        hs.append(h)
    o = h @ q["W4"].T + q["b4"]
    n = o.size
    loss = float(np.sum((o - yn) ** 2) / n)
    d = 2.0 * (o - yn) / n
    g = {"W4": d.T @ hs[3], "b4": d.sum(0)}
    d = (d @ q["W4"]) * dsilu(zs[2])
    g["W3"], g["b3"] = d.T @ hs[2], d.sum(0)
    d = (d @ q["W3"]) * dsilu(zs[1])
    g["W2"], g["b2"] = d.T @ hs[1], d.sum(0)
    d = (d @ q["W2"]) * dsilu(zs[0])
    g["W1"], g["b1"] = d.T @ hs[0], d.sum(0)
    flat = np.concatenate([g[name].ravel() for name, _ in SHAPES])
    return loss, flat

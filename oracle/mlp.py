"""fp64 CPU oracle of the learned-radius MLP -- TEST INFRASTRUCTURE.

The network has no reference counterpart; this restates its forward pass and
its gradient (mean squared error of the standardised target) in float64 with
the same bf16 rounding of weights and layer inputs that the GPU applies, so
the GPU gradient can be checked against it.  Its own derivative is pinned by
central finite differences in tests/test_mlp_oracle.py.

Parameter layout (float32 vector of 9488): W1[16][64], b1[64], W2[64][64],
b2[64], W3[64][64], b3[64], w4[64], b4, 15 unused.
"""
from __future__ import annotations

import numpy as np

NF, NH = 16, 64
P_W1, P_B1 = 0, 1024
P_W2, P_B2 = 1088, 5184
P_W3, P_B3 = 5248, 9344
P_W4, P_B4 = 9408, 9472
NPARAM = 9488


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return dict(W1=p[P_W1:P_B1].reshape(NF, NH), b1=p[P_B1:P_W2], W2=p[P_W2:P_B2].reshape(NH, NH),
                b2=p[P_B2:P_W3], W3=p[P_W3:P_B3].reshape(NH, NH), b3=p[P_B3:P_W4],
                w4=p[P_W4:P_B4], b4=p[P_B4])


def forward(p, norm, feats, rounding=True):
    """Returns (output on the standardised scale, cache)."""
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    w = unpack(p)
    f = np.asarray(feats, np.float64).reshape(-1, NF)
    x = r((f - norm[:NF]) * norm[NF:2 * NF])
    if rounding:
        W1, W2, W3 = bf16(w["W1"]), bf16(w["W2"]), bf16(w["W3"])
    else:
        W1, W2, W3 = w["W1"], w["W2"], w["W3"]
    z1 = x @ W1 + w["b1"]
    a1 = r(np.maximum(z1, 0))
    z2 = a1 @ W2 + w["b2"]
    a2 = r(np.maximum(z2, 0))
    z3 = a2 @ W3 + w["b3"]
    a3 = np.maximum(z3, 0)
    out = a3 @ w["w4"] + w["b4"]
    return out, dict(x=x, z1=z1, a1=a1, z2=z2, a2=a2, z3=z3, a3=a3, W1=W1, W2=W2, W3=W3, w=w)


def loss_and_grad(p, norm, feats, y, rounding=True):
    out, c = forward(p, norm, feats, rounding)
    t = (np.asarray(y, np.float64) - norm[2 * NF]) / norm[2 * NF + 1]
    n = len(t)
    e = out - t
    loss = float(np.mean(e * e))
    dout = 2.0 * e / n
    g = np.zeros(NPARAM, np.float64)
    w = c["w"]
    g[P_W4:P_B4] = c["a3"].T @ dout
    g[P_B4] = dout.sum()
    dz3 = np.outer(dout, w["w4"]) * (c["z3"] > 0)
    g[P_W3:P_B3] = (c["a2"].T @ dz3).ravel()
    g[P_B3:P_W4] = dz3.sum(0)
    dz2 = (dz3 @ c["W3"].T) * (c["z2"] > 0)
    g[P_W2:P_B2] = (c["a1"].T @ dz2).ravel()
    g[P_B2:P_W3] = dz2.sum(0)
    dz1 = (dz2 @ c["W2"].T) * (c["z1"] > 0)
    g[P_W1:P_B1] = (c["x"].T @ dz1).ravel()
    g[P_B1:P_W2] = dz1.sum(0)
    return loss, g

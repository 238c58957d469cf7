"""CPU pins of the MLP oracle (oracle/mlp.py): its analytic gradient equals
central finite differences of its own loss (float64, no bf16 rounding), and the
bf16 rounding matches the bit-level definition."""
import numpy as np

from oracle import mlp as M


def test_bf16_rounding_is_nearest_even():
    a = np.array([1.0, 1.00390625, 1.0078125, 1.01171875, -2.5, 3.0e-3], np.float32)
    r = M.bf16(a)
    # 1 + 2^-8 is halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even -> 1
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    # 1 + 3*2^-8 is halfway between 1+2^-7 and 1+2^-6: even mantissa -> 1 + 2^-6
    assert r[3] == 1.015625
    assert r[4] == -2.5
    u = np.array([r[5]], np.float32).view(np.uint32)[0]
    assert u & 0xFFFF == 0


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = np.zeros(M.NPARAM)
    p[M.P_W1:M.P_B1] = (rng.standard_normal((16, 64)) * 0.3).ravel()
    p[M.P_W1 + 12 * 64:M.P_B1] = 0.0
    p[M.P_B1:M.P_W2] = rng.standard_normal(64) * 0.1
    p[M.P_W2:M.P_B2] = (rng.standard_normal((64, 64)) * 0.15).ravel()
    p[M.P_B2:M.P_W3] = rng.standard_normal(64) * 0.1
    p[M.P_W3:M.P_B3] = (rng.standard_normal((64, 64)) * 0.15).ravel()
    p[M.P_B3:M.P_W4] = rng.standard_normal(64) * 0.1
    p[M.P_W4:M.P_B4] = rng.standard_normal(64) * 0.2
    p[M.P_B4] = 0.3
    norm = np.zeros(34)
    norm[:12] = rng.standard_normal(12)
    norm[16:28] = 1.0 / (0.5 + rng.random(12))
    norm[32], norm[33] = 2.0, 1.5
    f = np.zeros((40, 16))
    f[:, :12] = rng.standard_normal((40, 12))
    y = rng.standard_normal(40) * 2 + 1
    loss, g = M.loss_and_grad(p, norm, f, y, rounding=False)
    idx = np.concatenate([rng.choice(np.arange(0, 768), 15), rng.choice(np.arange(1024, 9473), 60),
                          [M.P_B4]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, norm, f, y, rounding=False)[0]
              - M.loss_and_grad(pm, norm, f, y, rounding=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])

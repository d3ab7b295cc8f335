"""Pins the CPU restatement of the indicator MLP (oracle/mlp_oracle.py), the
checker of the GPU training and inference kernels.

- backward: every gradient entry against central finite differences of the
  loss (float64), on random parameters and data, for both labels and a
  positive-class weight
- features: closed forms on piecewise-polynomial states
- Adam: the first step moves every parameter by -lr * sign(gradient)
"""
import numpy as np

from oracle import mlp_oracle as MO


def test_layout():
    assert MO.NPARAM == 8129
    p = np.arange(MO.NPARAM, dtype=np.float64)
    assert np.array_equal(MO.pack(*MO.unpack(p)), p)


def test_backward_vs_finite_differences():
    rng = np.random.default_rng(3)
    params = MO.init_params(seed=5) + 0.05 * rng.standard_normal(MO.NPARAM)
    X = rng.standard_normal((9, MO.NIN))
    Y = (rng.random(9) < 0.4).astype(np.float64)
    loss, g = MO.loss_grad(params, X, Y, pos_weight=3.0)
    # every bias and w3 entry, plus a random sample of weight entries
    probe = np.concatenate([np.arange(MO.OFF_B1, MO.OFF_W2), np.arange(MO.OFF_B2, MO.NPARAM),
                            rng.choice(MO.OFF_B1, 200, replace=False),
                            MO.OFF_W2 + rng.choice(MO.OFF_B2 - MO.OFF_W2, 200, replace=False)])
    h = 1e-6
    for k in probe:
        e = np.zeros(MO.NPARAM)
        e[k] = h
        fd = (MO.loss_grad(params + e, X, Y, 3.0)[0] - MO.loss_grad(params - e, X, Y, 3.0)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(fd), (k, fd, g[k])


def test_loss_closed_form():
    # all weights zero: p = sigmoid(b3) for every row
    p = np.zeros(MO.NPARAM)
    p[MO.OFF_B3] = 0.3
    X = np.ones((4, MO.NIN))
    Y = np.array([1.0, 0.0, 1.0, 0.0])
    loss, g = MO.loss_grad(p, X, Y, pos_weight=2.0)
    s = 1 / (1 + np.exp(-0.3))
    assert np.isclose(loss, (2 * 2 * -np.log(s) + 2 * -np.log(1 - s)) / 4)
    assert np.isclose(g[MO.OFF_B3], (2 * 2 * (s - 1) + 2 * s) / 4)


def _state(N=6):
    return np.zeros((N + 2, N + 2, 10, 8), order="F")


def test_features_closed_form():
    N, M, h = 6, 50.0, 0.1
    uh = _state(N)
    uh[:, :, 0, 0] = 2.0            # uniform density
    uh[:, :, 0, 4] = 5.0            # uniform energy
    uh[:, :, 0, 1] = -0.5           # uniform rho u < 0, rho v = 0
    cells = [(3, 4), (1, 1)]
    f = MO.features(uh, cells, M, h, h)
    expect = np.zeros(60)
    expect[0] = np.log10(M * h * h / (2.0 + 1e-12)) / 4
    expect[5] = np.log10(h) / 2
    for k in range(5):
        expect[12 * k + 10] = -0.5 / (0.5 + 1e-12)
    assert np.allclose(f, expect[None, :], rtol=0, atol=1e-15)
    # density linear in x: mean rho(i) = i, slope mode 1 = 0.5 in every cell
    for i in range(N + 2):
        uh[i, :, 0, 0] = float(i)
    uh[:, :, 1, 0] = 0.5
    f = MO.features(uh, [(3, 4)], M, h, h)[0]
    s = 4.0 + 1e-12                   # max |mean| over C(3), E(4), W(2), N(3), S(3)
    assert np.isclose(f[12 * 1 + 0], 1.0 / s) and np.isclose(f[12 * 2 + 0], -1.0 / s)   # E, W differences
    assert f[12 * 3 + 0] == 0.0 and f[12 * 4 + 0] == 0.0                                  # N, S differences
    for k in range(5):
        assert np.isclose(f[12 * k + 1], 0.5 / s)       # density x-slope
        assert f[12 * k + 2] == 0.0 and f[12 * k + 6] == 0.0


def test_adam_first_step():
    rng = np.random.default_rng(0)
    p0 = rng.standard_normal(50)
    g = rng.standard_normal(50)
    p1, m, v = MO.adam_step(p0, g, np.zeros(50), np.zeros(50), 1, lr=1e-3)
    assert np.allclose(p1 - p0, -1e-3 * np.sign(g) * np.abs(g) / (np.abs(g) + 1e-8), rtol=1e-12)

"""The surrogate network's CPU oracle, pinned by finite differences, and the parameter layout."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402


def _problem(n=12, seed=3):
    rng = np.random.default_rng(seed)
    p = np.zeros(mo.nparams())
    lay, (wo, bo), _ = mo.offsets()
    for l, (w, b) in enumerate(lay):
        fan = mo.IN if l == 0 else mo.HID
        W = rng.uniform(-1, 1, (mo.HID, mo.layer_in(l))) * np.sqrt(6.0 / fan)
        if l == 0:
            W[:, mo.IN:] = 0.0
        p[w:w + W.size] = W.ravel()
        p[b:b + mo.HID] = rng.normal(0, 0.1, mo.HID)
    p[wo:wo + mo.HID] = rng.normal(0, 0.1, mo.HID)
    p[bo] = 0.3
    par = rng.uniform(-1, 1, (3, mo.NPAR))
    sets = rng.integers(0, 3, n)
    q = rng.uniform(-1, 1, (n, 2))
    F = np.zeros((mo.NFREQ, 3))
    F[:, :2] = rng.normal(0, 2, (mo.NFREQ, 2))
    X = mo.features(par, sets, q, F, np.full(mo.NFREQ, -1), np.tile([0.0, 1.0], (mo.NPAR, 1)))
    t = rng.normal(0, 1, n)
    return p, X, t


def test_parameter_count_and_layout():
    # 70 inputs -> 4 x 256 -> 1: 215,809 parameters; the first layer is stored 128 wide
    assert mo.used_params() == 215809
    assert mo.nparams() == 215809 + mo.HID * (mo.IN_PAD - mo.IN)


def test_features_are_parameters_then_sin_cos():
    par = np.array([[0.1, -0.2, 0.3, -0.4, 0.5, -0.6]])
    F = np.zeros((mo.NFREQ, 3))
    F[0, :2] = (1.0, 0.0)
    F[1, 2] = 2.0                              # 2 |q| (offset + scale * par_3)
    pidx = np.full(mo.NFREQ, -1)
    pidx[1] = 3
    affine = np.tile([0.0, 1.0], (mo.NPAR, 1))
    affine[3] = (10.0, 5.0)                    # physical value 10 + 5 * (-0.4) = 8
    q = np.array([[0.25, 0.7]])
    X = mo.features(par, [0], q, F, pidx, affine)
    assert np.allclose(X[0, :6], par[0])
    assert np.isclose(X[0, 6], 1.0)          # sin(2 pi * 0.25)
    assert np.isclose(X[0, 6 + mo.NFREQ], 0.0, atol=1e-15)
    a1 = 2.0 * np.hypot(0.25, 0.7) * 8.0
    assert np.isclose(X[0, 7], np.sin(a1)) and np.isclose(X[0, 7 + mo.NFREQ], np.cos(a1))
    assert np.all(X[0, 8:6 + mo.NFREQ] == 0) and np.all(X[0, 8 + mo.NFREQ:mo.IN] == 1)
    assert np.all(X[0, mo.IN:] == 0)


def test_backward_matches_finite_differences():
    p, X, t = _problem()
    loss, g = mo.loss_and_grad(p, X, t)
    y, _ = mo.forward(p, X)
    assert np.isclose(loss, ((y - t) ** 2).mean())
    rng = np.random.default_rng(7)
    lay, (wo, bo), n = mo.offsets()
    idx = list(rng.integers(0, n, 60)) + [wo + 5, bo, lay[0][0] + 3, lay[3][1] + 17]
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_and_grad(pp, X, t)[0] - mo.loss_and_grad(pm, X, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])
    # padding columns of the first layer see zero inputs: zero gradient
    W0 = g[lay[0][0]:lay[0][0] + mo.HID * mo.IN_PAD].reshape(mo.HID, mo.IN_PAD)
    assert np.all(W0[:, mo.IN:] == 0)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -3.14159, 0.0])
    r = mo.to_bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0           # tie to even
    assert r[2] == 1.0 + 2 ** -7                 # below half an ulp (2^-8): rounds down
    assert abs(r[3] + 3.140625) < 1e-12
    assert r[4] == 0.0

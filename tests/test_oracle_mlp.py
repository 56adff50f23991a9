"""Pins the FP64 proposal-MLP oracle: analytic gradients against central finite differences of the
loss, softmax normalisation, GELU against scipy's erf form, and Adam against its closed form."""
import numpy as np
import pytest
from scipy import special


def params(rng, D=12, H=9, C=7, scale=0.4):
    return {"W1": rng.normal(scale=scale, size=(H, D)), "b1": rng.normal(scale=0.1, size=H),
            "W2": rng.normal(scale=scale, size=(H, H)), "b2": rng.normal(scale=0.1, size=H),
            "W3": rng.normal(scale=scale, size=(C, H)), "b3": rng.normal(scale=0.1, size=C)}


def test_forward_softmax_and_gelu(orc):
    rng = np.random.default_rng(0)
    P = params(rng)
    x = rng.normal(size=(5, 12))
    a = orc.mlp_forward(x, P)
    assert np.allclose(a["p"].sum(1), 1.0)
    gelu = lambda v: 0.5 * v * (1 + special.erf(v / np.sqrt(2)))
    h1 = gelu(x @ P["W1"].T + P["b1"])
    h2 = gelu(h1 @ P["W2"].T + P["b2"])
    z = h2 @ P["W3"].T + P["b3"]
    assert np.allclose(a["h1"], h1) and np.allclose(a["z"], z)
    assert np.allclose(a["p"], special.softmax(z, axis=1))


def test_gradients_match_finite_differences(orc):
    rng = np.random.default_rng(1)
    P = params(rng)
    x = rng.normal(size=(6, 12))
    t = rng.dirichlet(np.ones(7), size=6)
    a = orc.mlp_forward(x, P)
    g = orc.mlp_backward(x, t, P, a)
    eps = 1e-6
    for name in P:
        flat = P[name].reshape(-1)
        for k in rng.choice(flat.size, min(flat.size, 8), replace=False):
            old = flat[k]
            flat[k] = old + eps
            lp = orc.mlp_loss(orc.mlp_forward(x, P)["p"], t)
            flat[k] = old - eps
            lm = orc.mlp_loss(orc.mlp_forward(x, P)["p"], t)
            flat[k] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[name].reshape(-1)[k]) < 1e-6 * max(1.0, abs(fd)), (name, k)


def test_adam_closed_form_constant_gradient(orc):
    # for a constant gradient g the bias-corrected moments are exactly g and g^2, so every step
    # moves w by -lr * g / (|g| + eps)
    w = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -0.1, 2.0])
    m = np.zeros(3)
    v = np.zeros(3)
    for t in range(1, 6):
        orc.adam(w, g, m, v, t, lr=0.01)
    assert np.allclose(w, np.array([1.0, -2.0, 0.5]) - 5 * 0.01 * g / (np.abs(g) + 1e-8))

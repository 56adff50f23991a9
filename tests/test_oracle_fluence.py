"""Pins orc_fluence, the global fluence phi0 of the true-Poisson likelihood (the photon-unit
decision for a model the reference keeps at mean 1): phi0 = (mean photons per pattern over
unmasked pixels) / (mean slice sum over unmasked pixels with W > 0).

The checks are constructions whose answer is known without the oracle's loop: slices that are
constant on a chosen valid set and -1 elsewhere, patterns that are constant on the unmasked
pixels, with huge values planted on masked pixels and invalid samples so that any slip in the
two selections (mask, W > 0) or in the two denominators (patterns, rotations) moves the result.
"""
import numpy as np
import pytest


def _geometry(rng, P=64, n_masked=9):
    mask = np.ones(P, np.int32)
    mask[rng.choice(P, n_masked, replace=False)] = 0
    return mask


def test_constant_slices_and_patterns(orc):
    rng = np.random.default_rng(21)
    P, S, I = 64, 5, 3
    mask = _geometry(rng, P)
    unmasked = np.flatnonzero(mask)
    a = np.array([0.5, 1.25, 2.0, 3.5, 0.75])     # slice value on its valid set
    n = np.array([10, 30, 44, 1, 55])              # valid unmasked pixels of each slice
    b = np.array([2.0, 7.0, 0.5])                  # pattern value on every unmasked pixel
    W = np.full((S, P), 1.0e6)                      # masked pixels: large and "valid" (W > 0)
    for s in range(S):
        W[s, unmasked] = -1.0                        # invalid samples on unmasked pixels
        W[s, rng.choice(unmasked, n[s], replace=False)] = a[s]
    W[0, unmasked[W[0, unmasked] < 0][:3]] = 0.0     # W = 0 is not a valid sample either
    K = np.full((I, P), 9.0e5)                       # masked pixels: large counts that must not count
    for i in range(I):
        K[i, unmasked] = b[i]
    expected = (b.mean() * unmasked.size) / (a * n).mean()
    assert orc.fluence(W, K, mask) == pytest.approx(expected, rel=1e-13)


def test_pattern_proportional_to_one_slice_gives_the_factor(orc):
    """One rotation, one pattern K = c W on the valid pixels and 0 on the invalid ones: phi0 = c."""
    rng = np.random.default_rng(22)
    P = 80
    mask = _geometry(rng, P)
    W = rng.uniform(0.1, 3.0, size=(1, P))
    W[0, rng.choice(P, 12, replace=False)] = -1.0
    c = 37.5
    K = np.where((W > 0) & (mask[None, :] != 0), c * W, 0.0)
    K[0, mask == 0] = 1.0e4
    assert orc.fluence(W, K, mask) == pytest.approx(c, rel=1e-13)


def test_scaling_laws_and_replication(orc):
    rng = np.random.default_rng(23)
    P, S, I = 50, 7, 4
    mask = _geometry(rng, P)
    W = rng.uniform(0.0, 2.0, size=(S, P))
    W[rng.random((S, P)) < 0.2] = -1.0
    K = rng.poisson(3.0, size=(I, P)).astype(np.float64)
    f = orc.fluence(W, K, mask)
    assert orc.fluence(W, 4.0 * K, mask) == pytest.approx(4.0 * f, rel=1e-13)        # linear in K
    assert orc.fluence(2.5 * W, K, mask) == pytest.approx(f / 2.5, rel=1e-13)        # 1/W (negatives stay negative)
    assert orc.fluence(np.vstack([W, W]), np.vstack([K, K, K]), mask) == pytest.approx(f, rel=1e-13)  # means
    # brute force over the definition's two sets, in pure Python
    sk = sum(K[i, p] for i in range(I) for p in range(P) if mask[p])
    sw = sum(W[s, p] for s in range(S) for p in range(P) if mask[p] and W[s, p] > 0)
    assert f == pytest.approx((sk / I) / (sw / S), rel=1e-12)


def test_fallbacks(orc):
    P = 16
    mask = np.ones(P, np.int32)
    W = np.full((3, P), -1.0)
    K = np.ones((2, P))
    assert orc.fluence(W, K, mask) == 1.0               # no valid sample
    assert orc.fluence(np.ones((3, P)), np.zeros((2, P)), mask) == 1.0  # no photon
    assert orc.fluence(np.ones((3, P)), K, np.zeros(P, np.int32)) == 1.0  # every pixel masked

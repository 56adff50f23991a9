"""Model initialisation modes and pattern normalisation (emc_helper.cpp:371-534): the host library
against the oracle, and the oracle against closed forms and brute force.  CPU only."""
import numpy as np
import pytest

import emc_amd

L = 12


def shells_numpy(L, K):
    """Brute-force shell average of the patterns (integer pixel radius below L/2)."""
    c = np.arange(L) - L / 2 + 0.5
    r = np.sqrt(c[None, :] ** 2 + c[:, None] ** 2)  # [y][x]
    sh = np.floor(r).astype(int)
    ok = r < L / 2
    avg = np.zeros(L // 2 + 1)
    for s in range(L // 2):
        sel = ok & (sh == s)
        vals = K.reshape(-1, L, L)[:, sel].astype(np.float64)
        avg[s] = vals.mean() if vals.size else 0.0
    return avg


def test_radial_model_constant_and_brute_force(orc):
    rng = np.random.default_rng(1)
    Kc = np.full((3, L * L), 2.5, np.float32)
    for M in (orc.model_radial(L, Kc), emc_amd.model_init_radial(L, Kc)):
        c = np.arange(L) - L / 2 + 0.5
        rad = np.sqrt(c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2).ravel()
        inside = rad < L / 2
        outer = inside & (rad >= L / 2 - 1)  # interpolates towards the zero shell past L/2
        assert np.allclose(M[inside & ~outer], 2.5) and np.all(M[~inside] == -1)
        assert np.allclose(M[outer], 2.5 * (1 - (rad[outer] - np.floor(rad[outer]))))
    K = rng.poisson(3.0, size=(5, L * L)).astype(np.float32)
    Mo = orc.model_radial(L, K)
    avg = shells_numpy(L, K)
    c = np.arange(L) - L / 2 + 0.5
    rad = np.sqrt(c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2)  # [z][y][x]
    r0 = np.floor(rad).astype(int)
    want = np.where(rad < L / 2, avg[np.minimum(r0, L // 2)] * (1 - (rad - r0)) +
                    avg[np.minimum(r0 + 1, L // 2)] * (rad - r0), -1.0)
    assert np.allclose(Mo, want.ravel(), rtol=1e-12, atol=1e-12)
    assert np.allclose(emc_amd.model_init_radial(L, K), Mo, rtol=1e-6, atol=1e-6)


def test_orientation_model_identity_and_library(orc):
    """One pattern at the identity rotation lands on plane z = L/2 + 1, shifted by (+1, +1): the
    host insert's round(L/2 + 0.5 + r) (emc_helper.cpp:115-117)."""
    coords = emc_amd.coords_flat(L)
    mask = np.ones(L * L, np.int32)
    mask[:3] = 0
    rng = np.random.default_rng(2)
    K = rng.uniform(1, 5, size=(1, L * L)).astype(np.float32)
    M = orc.model_orientations(L, K, mask, coords, np.array([[1, 0, 0, 0]], np.float32)).reshape(L, L, L)
    img = (K[0] * (mask == 1)).reshape(L, L)  # [y][x]
    want = np.zeros((L, L))
    want[1:, 1:] = img[:-1, :-1]
    assert np.allclose(M[L // 2 + 1], want)
    assert np.count_nonzero(M) == np.count_nonzero(want)
    # the library with its own seeded orientations equals the oracle given those orientations
    Ks = rng.poisson(2.0, size=(40, L * L)).astype(np.float32)
    m, q = emc_amd.model_init_orientations(L, Ks, mask, coords, seed=9)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-6)
    assert np.allclose(m, orc.model_orientations(L, Ks, mask, coords, q), rtol=1e-6, atol=1e-6)
    m2, q2 = emc_amd.model_init_orientations(L, Ks, mask, coords, seed=9)
    assert np.array_equal(q, q2) and np.array_equal(m, m2)  # seeded


def test_normalize_patterns(orc):
    rng = np.random.default_rng(3)
    P = L * L
    mask = (rng.random(P) < 0.8).astype(np.int32)
    K = rng.poisson(4.0, size=(6, P)).astype(np.float32)
    K[2] = 0.0  # a zero sum leaves the pattern as it is
    out = emc_amd.normalize_patterns(K, mask)
    ref = orc.normalize_patterns(K, mask)
    assert np.allclose(out, ref, rtol=1e-6)
    s = (ref * (mask == 1)).sum(1)
    assert np.allclose(np.delete(s, 2), P) and np.all(ref[2] == 0)


def test_model_read_raw(tmp_path):
    rng = np.random.default_rng(4)
    m = rng.normal(size=L ** 3).astype(np.float32)
    p = tmp_path / "model.bin"
    m.tofile(p)
    assert np.array_equal(emc_amd.model_read_raw(p, L), np.abs(m))
    m[:-1].tofile(p)
    with pytest.raises(emc_amd.EmcError):
        emc_amd.model_read_raw(p, L)
    with pytest.raises(emc_amd.EmcError):
        emc_amd.model_read_raw(tmp_path / "missing.bin", L)

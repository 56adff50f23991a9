"""Host-side product code (no GPU): libemc's rotation sampling, geometry and synthetic data
generator against the oracle, closed forms and numpy; and that libemc exports every symbol
include/emc.h declares."""
import os
import re

import numpy as np
import pytest

import emc_amd
from emc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "emc.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(emc_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(L, s), s
        assert s in _lib.SIGNATURES, f"{s} has no ctypes signature"


@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8])
def test_rotations_bit_exact_vs_oracle(orc, n):
    q, w = emc_amd.rotations(n)
    q0, w0, _ = orc.rotations(n)
    assert np.array_equal(q, q0) and np.array_equal(w, w0)


def test_rotation_count_errors():
    assert emc_amd.rotation_count(0) < 0
    with pytest.raises(ValueError):
        emc_amd.rotations(0)


def test_coords_and_mask_bit_exact_vs_oracle(orc):
    for L in (16, 64):
        assert all(np.array_equal(a, b) for a, b in zip(emc_amd.coords_flat(L), orc.coords_flat(L)))
        e1 = emc_amd.coords_ewald(L, 3e-4, 0.74, 1.03e-9)
        e2 = orc.coords_ewald(L, 3e-4, 0.74, 1.03e-9)
        assert all(np.array_equal(a, b) for a, b in zip(e1, e2))
        assert np.array_equal(emc_amd.mask_annulus(L, 4, L / 2), orc.mask_annulus(L, 4, L / 2))


def test_sigma_schedule_matches_oracle(orc):
    for it in range(0, 60, 7):
        assert emc_amd.sigma_schedule(it, 0.4, 0.12, 25) == orc.sigma_at(it, 0.4, 0.12, 25)


def test_synth_model_is_shifted_power_spectrum_of_object():
    L = 24
    obj = emc_amd.synth_object(L, 30, 5.0, 7).reshape(L, L, L).astype(np.float64)
    model = emc_amd.synth_model(L, 30, 5.0, 7).reshape(L, L, L)
    ref = np.fft.fftshift(np.abs(np.fft.fftn(obj)) ** 2)
    ref /= ref.mean()
    assert np.allclose(model, ref, rtol=2e-5, atol=1e-6 * ref.max())
    assert abs(model.mean() - 1.0) < 1e-5
    # centrosymmetry of |F|^2 about the zero frequency at index L/2
    flip = np.roll(model[::-1, ::-1, ::-1], 1, axis=(0, 1, 2))
    assert np.allclose(model, flip, rtol=1e-4, atol=1e-6 * model.max())


def test_synth_object_atoms_inside_sphere():
    L, R = 32, 6.0
    obj = emc_amd.synth_object(L, 40, R, 3).reshape(L, L, L)
    z, y, x = np.nonzero(obj > 0.5)
    r = np.sqrt((x - L / 2) ** 2 + (y - L / 2) ** 2 + (z - L / 2) ** 2)
    assert r.max() < R + 1.5
    assert abs(obj.sum() - 40 * (2 * np.pi) ** 1.5) < 0.01 * 40 * (2 * np.pi) ** 1.5


def test_synth_patterns_statistics_and_geometry(orc):
    L = 32
    model = emc_amd.synth_model(L, 40, 6.0, 1)
    coords = emc_amd.coords_flat(L)
    mask = emc_amd.mask_annulus(L, 3, 16)
    K, q = emc_amd.synth_patterns(model, L, coords, mask, 400, 2000.0, 5)
    K2, q2 = emc_amd.synth_patterns(model, L, coords, mask, 400, 2000.0, 5)
    assert np.array_equal(K, K2) and np.array_equal(q, q2)     # seeded, thread-count independent
    assert np.all(K[:, mask == 0] == 0) and K.min() >= 0
    tot = K.sum(1)
    assert abs(tot.mean() - 2000) < 3 * np.sqrt(2000 / 400) + 1
    assert np.allclose(np.linalg.norm(q, axis=1), 1, atol=1e-6)
    W = orc.slices(model.astype(np.float64), L, q[:50], coords, "trilinear")
    lam = np.where(mask[None] == 1, np.maximum(W, 0), 0)
    lam *= 2000.0 / lam.sum(1, keepdims=True)
    # expected counts follow the oracle slice at the true orientation
    resid = (K[:50] - lam)[:, mask == 1]
    assert abs(resid.mean()) < 0.05
    assert abs((resid ** 2 / np.maximum(lam[:, mask == 1], 1e-9)).mean() - 1.0) < 0.1  # Poisson variance


def test_rotation_distribution_is_uniform():
    L = 8
    model = np.ones(L ** 3, np.float32)
    coords = emc_amd.coords_flat(L)
    mask = np.ones(L * L, np.int32)
    _, q = emc_amd.synth_patterns(model, L, coords, mask, 20000, 10.0, 9)
    # Haar measure: each component squared has mean 1/4; a component has density
    # (2/pi) sqrt(1 - x^2) on [-1, 1], so the mean of its absolute value is 4/(3 pi)
    assert np.allclose((q.astype(np.float64) ** 2).mean(0), 0.25, atol=0.01)
    assert abs(np.abs(q[:, 0]).mean() - 4 / (3 * np.pi)) < 0.01


@pytest.mark.parametrize("bad", [dict(mfma_staging=2), dict(keep_logits=-1), dict(keep_logits=2),
                                 dict(tile_rows=-1), dict(count_bytes=1), dict(mstep=3)])
def test_create_rejects_bad_option_values(bad):
    """emc_create checks every option before it touches a device (so this runs without a GPU)."""
    with pytest.raises(emc_amd.EmcError) as ei:
        emc_amd.Engine(32, 60, 10, **bad)
    assert "bad option value" in str(ei.value)

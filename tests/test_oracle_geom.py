"""Pins the oracle's geometry: detector coordinates (emc_helper.cpp:55-79), masks, rotation
matrices (emc_cuda.cu:113-121) and the sigma schedule (main.cpp:251) against closed forms and
scipy's rotation routine."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation


def test_flat_coords_closed_form(orc):
    L = 16
    x, y, z = orc.coords_flat(L)
    xs, ys = np.meshgrid(np.arange(L), np.arange(L))  # ys varies along rows -> index y*L + x
    assert np.array_equal(x, (xs - L // 2 + 0.5).ravel().astype(np.float32))
    assert np.array_equal(y, (ys - L // 2 + 0.5).ravel().astype(np.float32))
    assert np.all(z == 0)


def test_ewald_coords_closed_form(orc):
    L, pix, D, lam = 64, 75e-6 * 4, 0.74, 1.03e-9
    x, y, z = orc.coords_ewald(L, pix, D, lam)
    xs, ys = np.meshgrid(np.arange(L), np.arange(L))
    X = (xs - L // 2 + 0.5).ravel()
    Y = (ys - L // 2 + 0.5).ravel()
    r = np.hypot(X, Y)
    t = np.arctan2(r * pix, D)
    Z = r * (1 - np.cos(t)) / np.sin(t)
    assert np.array_equal(x, X.astype(np.float32)) and np.array_equal(y, Y.astype(np.float32))
    assert np.allclose(z, Z, rtol=1e-5, atol=1e-9)
    # curvature is small and positive, Z ~ r^2 * pix / (2 D) for small angles
    assert np.allclose(z, r ** 2 * pix / (2 * D), rtol=1e-3)


@pytest.mark.parametrize("L,rb,ro,count", [(32, 3, 16, 780), (64, 4, 32, 3176), (128, 4, 64, 12840)])
def test_annulus_mask_counts(orc, L, rb, ro, count):
    m = orc.mask_annulus(L, rb, ro)
    assert m.sum() == count
    xs, ys = np.meshgrid(np.arange(L), np.arange(L))
    r = np.hypot(xs - L / 2 + 0.5, ys - L / 2 + 0.5).ravel()
    assert np.array_equal(m, ((r >= rb) & (r <= ro)).astype(np.int32))


def test_rotmat_matches_scipy(orc):
    rng = np.random.default_rng(0)
    for _ in range(50):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        m = orc.rotmat(q.astype(np.float32))
        ref = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()  # scipy is scalar-last
        assert np.allclose(m, ref, atol=2e-6)


def test_sigma_schedule(orc):
    # sigma_final + (sigma_start - sigma_final) * 2^(-it / half_life), stored as float
    for it in [0, 1, 7, 25, 100]:
        v = orc.sigma_at(it, 0.4, 0.12, 25)
        assert abs(v - (0.12 + 0.28 * 2 ** (-it / 25))) < 1e-7

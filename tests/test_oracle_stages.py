"""Pins the oracle's iteration stages against closed forms, invariants and independent library
routines (scipy.ndimage trilinear interpolation, scipy.stats Poisson pmf, scipy.special
logsumexp, numpy FFT, numpy matmul)."""
import numpy as np
import pytest
from scipy import ndimage, special, stats

L = 16


def rand_model(rng, L=L, lo=0.5, hi=2.0):
    return rng.uniform(lo, hi, size=L ** 3)


def rand_quats(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[q[:, 0] < 0] *= -1
    return q.astype(np.float32)


def test_identity_rotation_reads_a_model_plane(orc):
    rng = np.random.default_rng(1)
    M = rand_model(rng)
    coords = orc.coords_flat(L)
    q = np.array([[1, 0, 0, 0]], np.float32)
    Wn = orc.slices(M, L, q, coords, "nearest")[0].reshape(L, L)
    Wt = orc.slices(M, L, q, coords, "trilinear")[0].reshape(L, L)
    vol = M.reshape(L, L, L)  # [z][y][x]
    # u = (L/2 - 0.5) + X = x exactly; z: lroundf(L/2 - 0.5) = L/2 (half away from zero)
    plane = vol[L // 2]
    assert np.array_equal(Wn[:, 1:][1:], plane[1:, 1:])      # nearest excludes index 0 (bounds > 0)
    assert np.all(Wn[0, :] == -1) and np.all(Wn[:, 0] == -1)
    # trilinear at z = L/2 - 0.5: mean of the two planes
    assert np.allclose(Wt, 0.5 * (vol[L // 2 - 1] + vol[L // 2]), rtol=1e-12)


def test_quarter_turn_about_z_permutes_pixels(orc):
    rng = np.random.default_rng(2)
    M = rand_model(rng)
    coords = orc.coords_flat(L)
    c = np.float32(np.sqrt(0.5))
    q = np.array([[c, 0, 0, c]], np.float32)  # +90 deg about z: (X, Y) -> (-Y, X)
    W = orc.slices(M, L, q, coords, "nearest")[0].reshape(L, L)  # W[y][x]
    vol = M.reshape(L, L, L)
    for y in range(2, L - 1):
        for x in range(2, L - 1):
            vx, vy = L - 1 - y, x
            assert W[y, x] == vol[L // 2, vy, vx]


def test_trilinear_matches_scipy_map_coordinates(orc):
    rng = np.random.default_rng(3)
    M = rand_model(rng)
    coords = orc.coords_flat(L)
    q = rand_quats(rng, 4)
    W = orc.slices(M, L, q, coords, "trilinear")
    X, Y, Z = (c.astype(np.float64) for c in coords)
    vol = M.reshape(L, L, L)
    for s in range(4):
        m = orc.rotmat(q[s]).astype(np.float64)
        u = (L / 2 - 0.5) + m @ np.stack([X, Y, Z])
        inside = np.all((u >= 1e-3) & (u <= L - 1 - 1e-3), axis=0)
        ref = ndimage.map_coordinates(vol, [u[2], u[1], u[0]], order=1, mode="nearest")
        assert inside.sum() > 50
        assert np.allclose(W[s][inside], ref[inside], rtol=1e-5)
        outside = np.any((u < -1e-3) | (u > L - 1 + 1e-3), axis=0)
        assert np.all(W[s][outside] == -1)


def test_trilinear_insert_is_adjoint_of_slice(orc):
    rng = np.random.default_rng(4)
    M = rand_model(rng)
    coords = orc.coords_flat(L)
    mask = np.ones(L * L, np.int32)
    q = rand_quats(rng, 3)
    X = rng.uniform(0, 5, size=(1, L * L))
    W = orc.slices(M, L, q, coords, "trilinear")
    for s in range(3):
        P = np.zeros((3, 1))
        P[s, 0] = 1.0
        Mnum, Wt, T = orc.mstep(P, X, mask, q, coords, L, "trilinear")
        valid = W[s] > 0
        lhs = (W[s][valid] * X[0][valid]).sum()
        rhs = (M * Mnum).sum()
        assert abs(lhs - rhs) < 1e-9 * abs(lhs)
        assert abs(Wt.sum() - valid.sum()) < 1e-9 * valid.sum()  # weights of a splat sum to 1


def test_true_poisson_equals_scipy_logpmf(orc):
    rng = np.random.default_rng(5)
    S, I, P = 5, 4, 40
    W = rng.uniform(0.1, 3.0, size=(S, P))
    W[0, :3] = -1.0  # invalid samples are excluded
    K = rng.poisson(2.0, size=(I, P)).astype(np.float64)
    mask = np.ones(P, np.int32)
    mask[5:9] = 0
    w = rng.uniform(0.5, 1.5, size=S)
    phi0 = 1.7
    R, ns = orc.logl(W, K, mask, w, "true_poisson", phi0=phi0)
    for s in range(S):
        v = (mask != 0) & (W[s] > 0)
        assert ns[s] == v.sum()
        for i in range(I):
            ref = stats.poisson.logpmf(K[i, v], phi0 * W[s, v]).sum() + special.gammaln(K[i, v] + 1).sum()
            assert abs(R[s, i] - (np.log(w[s]) + ref)) < 1e-9 * max(1, abs(ref))


def test_true_poisson_gemm_identity_and_constant_slice(orc):
    rng = np.random.default_rng(6)
    S, I, P = 7, 9, 33
    W = rng.uniform(0.2, 2.0, size=(S, P))
    K = rng.poisson(1.5, size=(I, P)).astype(np.float64)
    mask = np.ones(P, np.int32)
    w = np.full(S, 1.0 / S)
    R, _ = orc.logl(W, K, mask, w, "true_poisson")
    G = np.log(W) @ K.T - W.sum(1, keepdims=True) + np.log(w)[:, None]
    assert np.allclose(R, G, rtol=1e-12, atol=1e-10)
    Wc = np.full((1, P), 0.37)
    Rc, _ = orc.logl(Wc, K, mask, np.ones(1), "true_poisson")
    assert np.allclose(Rc[0], K.sum(1) * np.log(0.37) - P * 0.37)


@pytest.mark.parametrize("kind", ["poisson", "absolute", "relative"])
def test_chi2_kinds_closed_forms(orc, kind):
    rng = np.random.default_rng(7)
    S, P, sigma = 3, 20, 0.3
    W = rng.uniform(0.5, 2.0, size=(S, P))
    delta = 0.25
    K = np.stack([W[0], W[0] + delta])  # pattern 0 equals slice 0; pattern 1 is offset
    mask = np.ones(P, np.int32)
    w = np.array([0.2, 0.3, 0.5])
    R, ns = orc.logl(W, K, mask, w, kind, sigma=sigma)
    assert abs(R[0, 0] - np.log(0.2)) < 1e-12           # perfect match: D = 0
    if kind == "poisson":
        D = (delta ** 2 / (W[0] + 0.02)).sum()
    elif kind == "absolute":
        D = P * delta ** 2
    else:
        D = ((delta / W[0]) ** 2).sum()
    assert abs(R[0, 1] - (np.log(0.2) - D / (2 * P * sigma ** 2))) < 1e-10


def test_scaling_closed_forms_and_brute_force(orc):
    """phi[s, i] = sum K^2 / sum K W over unmasked pixels with W, K > 1e-10 (emc_cuda.cu:438-455)."""
    rng = np.random.default_rng(17)
    S, P = 3, 24
    W = rng.uniform(0.1, 3.0, size=(S, P))
    W[1, :5] = 0.0  # excluded pixels of slice 1
    mask = np.ones(P, np.int32)
    mask[-3:] = 0
    c = np.array([0.5, 2.0, 7.25])
    K = c[:, None] * W[0][None]  # pattern i is c_i times slice 0
    K[2, 7] = 0.0  # a zero count drops its pixel from both sums
    phi = orc.scaling(W, K, mask)
    assert np.allclose(phi[0], c, rtol=1e-13)
    one = np.zeros((1, P))
    one[0, 4] = 3.0  # one counted pixel: phi = K/W there
    assert np.isclose(orc.scaling(W, one, mask)[0, 0], 3.0 / W[0, 4])
    assert orc.scaling(W, one, mask)[1, 0] == 1.0  # W = 0 there: no pixel, phi := 1
    Kr = rng.poisson(1.5, size=(5, P)).astype(np.float64)
    phr = orc.scaling(W, Kr, mask)
    for s in range(S):
        for i in range(5):
            num = den = 0.0
            for p in range(P):
                if mask[p] and W[s, p] > 1e-10 and Kr[i, p] > 1e-10:
                    num += Kr[i, p] ** 2
                    den += Kr[i, p] * W[s, p]
            assert abs(phr[s, i] - (num / den if num > 0 and den > 0 else 1.0)) < 1e-12


def test_scaling_reference_mode_truncates(orc):
    """The reference's int accumulation of sum K W: exact when every product and partial sum is an
    integer, below the real sum (phi above) otherwise."""
    rng = np.random.default_rng(19)
    S, P = 3, 1500
    mask = np.ones(P, np.int32)
    Wi = rng.integers(1, 4, size=(S, P)).astype(np.float64)
    K = rng.poisson(1.5, size=(4, P)).astype(np.float64)
    assert np.allclose(orc.scaling_reference(Wi, K, mask), orc.scaling(Wi, K, mask), rtol=1e-6)
    Wf = rng.uniform(0.2, 2.0, size=(S, P))
    ref, exact = orc.scaling_reference(Wf, K, mask), orc.scaling(Wf, K, mask)
    assert np.all(ref > exact * 1.05)  # truncation drops a large share of sum K W at these counts
    zero = orc.scaling_reference(Wf, np.zeros((1, P)), mask)
    assert not np.isfinite(zero).any()  # 0 / 0, as the reference


@pytest.mark.parametrize("kind", ["poisson", "absolute"])
def test_scaled_copy_of_a_slice_is_a_perfect_match(orc, kind):
    """With phi from the scaling step, a pattern that is c times slice s has K/phi = W: D = 0."""
    rng = np.random.default_rng(18)
    S, P = 4, 30
    W = rng.uniform(0.5, 2.0, size=(S, P))
    mask = np.ones(P, np.int32)
    K = np.stack([3.0 * W[2], 0.4 * W[1]])
    w = np.full(S, 0.25)
    phi = orc.scaling(W, K, mask)
    R, _ = orc.logl(W, K, mask, w, kind, sigma=0.2, phi=phi)
    assert abs(R[2, 0] - np.log(0.25)) < 1e-12 and abs(R[1, 1] - np.log(0.25)) < 1e-12
    assert R[2, 0] == R[:, 0].max() and R[1, 1] == R[:, 1].max()
    R1, _ = orc.logl(W, K, mask, w, kind, sigma=0.2)  # unscaled, the copies no longer match
    assert R1[2, 0] < np.log(0.25) - 1.0


def test_softmax_matches_logsumexp(orc):
    rng = np.random.default_rng(8)
    R = rng.normal(scale=50.0, size=(300, 17)) - 1e4
    Pm, lse, tot = orc.softmax(R)
    ref = special.logsumexp(R, axis=0)
    assert np.allclose(lse, ref, rtol=0, atol=1e-9)
    assert np.allclose(Pm, np.exp(R - ref), atol=1e-15)
    assert np.allclose(Pm.sum(0), 1.0)
    assert abs(tot - ref.sum()) < 1e-7
    P2, lse2, _ = orc.softmax(R + 123.0)  # shift invariance
    assert np.allclose(P2, Pm) and np.allclose(lse2, lse + 123.0)


def test_mstep_one_hot_reinserts_pattern_and_conserves(orc):
    rng = np.random.default_rng(9)
    coords = orc.coords_flat(L)
    mask = np.ones(L * L, np.int32)
    mask[:L] = 0
    q = np.array([[1, 0, 0, 0]], np.float32)
    K = rng.poisson(3.0, size=(1, L * L)).astype(np.float64)
    Mnum, Wt, T = orc.mstep(np.ones((1, 1)), K, mask, q, coords, L, "nearest")
    vol = Mnum.reshape(L, L, L)
    Kimg = (K[0] * mask).reshape(L, L)
    assert np.allclose(vol[L // 2], Kimg)  # insertion bounds include index 0 (emc_atomic.cu:67)
    assert np.allclose(Wt.reshape(L, L, L)[L // 2], mask.reshape(L, L))
    assert T[0] == 1.0
    Mh, _, _ = orc.mstep(np.ones((1, 1)), K, mask, q, coords, L, "nearest", phi=np.full((1, 1), 4.0))
    assert np.allclose(Mh, Mnum / 4.0)  # U sums K P / phi (emc_atomic.cu:107-108)
    # general P: sum Mnum = sum_s sum_p U[s,p], sum Wt = sum_s T_s * n_valid_s (trilinear)
    S, I = 4, 6
    q = rand_quats(rng, S)
    Pm = rng.dirichlet(np.ones(S), size=I).T
    K = rng.poisson(2.0, size=(I, L * L)).astype(np.float64)
    Mnum, Wt, T = orc.mstep(Pm, K, mask, q, coords, L, "trilinear")
    U = Pm @ K
    assert np.allclose(T, Pm.sum(1))
    W = orc.slices(np.ones(L ** 3), L, q, coords, "trilinear")  # validity of each (s, p)
    valid = (W > 0) & (mask[None] == 1)
    assert np.isclose(Mnum.sum(), (U * valid).sum(), rtol=1e-12)
    assert np.isclose(Wt.sum(), (T[:, None] * valid).sum(), rtol=1e-12)


def test_mstep_skips_tiny_slices(orc):
    coords = orc.coords_flat(L)
    mask = np.ones(L * L, np.int32)
    q = np.array([[1, 0, 0, 0]], np.float32)
    Mnum, Wt, T = orc.mstep(np.full((1, 1), 1e-11), np.ones((1, L * L)), mask, q, coords, L, "nearest")
    assert Mnum.sum() == 0 and Wt.sum() == 0  # T_s <= 1e-10 is not inserted (emc_atomic.cu:190)


def test_compress_divide(orc):
    Mnum = np.array([1.0, 2.0, 3.0, 0.0])
    Wt = np.array([2.0, 0.0, 1.5, -1.0])
    assert np.array_equal(orc.compress(Mnum, Wt), [0.5, -1.0, 2.0, -1.0])


def test_blur_equals_fft_multiplier(orc):
    rng = np.random.default_rng(10)
    Lb, sigma = 12, 0.8
    M = rng.uniform(0, 2, size=Lb ** 3)
    M[rng.choice(Lb ** 3, 40, replace=False)] = -1.0
    out = orc.blur(M, Lb, sigma)
    k = np.arange(Lb)
    d = np.minimum(k, Lb - k)
    h = np.exp(-2 * np.pi ** 2 * sigma ** 2 * d ** 2 / Lb ** 2)
    H = h[:, None, None] * h[None, :, None] * h[None, None, :]
    A = np.where(M < 0, 0, M).reshape(Lb, Lb, Lb)
    ref = np.real(np.fft.ifftn(np.fft.fftn(A) * H)).ravel()
    keep = M >= 0
    assert np.allclose(out[keep], ref[keep], atol=1e-12)
    assert np.all(out[~keep] == -1)  # a missing voxel keeps its value
    Ms = np.where(M < 0, -0.25, M)    # ... also after normalisation scaled it (-1/avg)
    assert np.all(orc.blur(Ms, Lb, sigma)[~keep] == -0.25)
    c = orc.blur(np.full(Lb ** 3, 2.5), Lb, sigma)  # constants are preserved (h(0) = 1)
    assert np.allclose(c, 2.5)
    g = orc.blur_taps(Lb, 1e-9)  # sigma -> 0 is the identity
    assert np.allclose(g, np.eye(Lb)[0], atol=1e-12)


def test_normalize_mean_of_positive_is_one(orc):
    M = np.array([2.0, 4.0, -1.0, 0.0, 6.0])
    out, avg, cnt = orc.normalize(M)
    assert cnt == 3 and avg == 4.0
    assert np.allclose(out, M / 4.0)  # the -1 voxel is scaled too (emc_cuda.cu:795)
    assert np.isclose(out[out > 0].mean(), 1.0)


def test_normalize_over_the_weighted_support(orc):
    """The iteration's normaliser: the mean over the voxels with weight, zeros included; equal to
    the reference's mean over positive voxels when no supported voxel is zero; missing voxels
    (no weight, -1) are scaled but not averaged."""
    M = np.array([2.0, 4.0, -1.0, 0.0, 6.0, 0.0])
    Wt = np.array([1.0, 0.5, 0.0, 2.0, 3.0, 0.0])
    out, avg, cnt = orc.normalize_support(M, Wt)
    assert cnt == 4 and avg == 3.0                   # (2 + 4 + 0 + 6) / 4
    assert np.allclose(out, M / 3.0)
    M2 = np.array([2.0, 4.0, -1.0, 6.0])
    W2 = np.array([1.0, 1.0, 0.0, 1.0])
    assert np.allclose(orc.normalize_support(M2, W2)[0], orc.normalize(M2)[0])
    z, a0, c0 = orc.normalize_support(np.zeros(3), np.ones(3))  # nothing to normalise: unchanged
    assert c0 == 0 and a0 == 0.0 and np.all(z == 0.0)


def test_iterate_is_the_composition_of_stages(orc):
    rng = np.random.default_rng(11)
    Ls = 12
    coords = orc.coords_flat(Ls)
    mask = orc.mask_annulus(Ls, 1, 6)
    q, w, _ = orc.rotations(2)
    M = rng.uniform(0.2, 1.0, size=Ls ** 3)
    K = rng.poisson(1.0, size=(20, Ls * Ls)).astype(np.float64) * mask
    out = orc.iterate(M, Ls, q, w.astype(np.float64), K, mask, coords, "trilinear", "true_poisson",
                      blur_sigma=0.5, want_P=True)
    W = orc.slices(M, Ls, q, coords, "trilinear")
    phi0 = orc.fluence(W, K, mask)
    R, _ = orc.logl(W, K, mask, w.astype(np.float64), "true_poisson", phi0=phi0)
    Pm, lse, tot = orc.softmax(R)
    Mnum, Wt, T = orc.mstep(Pm, K, mask, q, coords, Ls, "trilinear")
    Mc, avg, _ = orc.normalize_support(orc.compress(Mnum, Wt), Wt)  # divide, normalise, then blur
    Mc = orc.blur(Mc, Ls, 0.5)
    assert np.allclose(out["P"], Pm) and np.allclose(out["lse"], lse)
    assert np.allclose(out["model"], Mc) and np.isclose(out["total"], tot)
    assert np.isclose(out["phi0"], phi0)


def test_iterate_with_scaling_is_the_composition_of_stages(orc):
    rng = np.random.default_rng(12)
    Ls = 12
    coords = orc.coords_flat(Ls)
    mask = orc.mask_annulus(Ls, 1, 6)
    q, w, _ = orc.rotations(2)
    M = rng.uniform(0.2, 1.0, size=Ls ** 3)
    K = rng.poisson(1.0, size=(20, Ls * Ls)).astype(np.float64) * mask
    out = orc.iterate(M, Ls, q, w.astype(np.float64), K, mask, coords, "trilinear", "poisson",
                      sigma=0.4, want_P=True, scaling=True)
    W = orc.slices(M, Ls, q, coords, "trilinear")
    phi = orc.scaling(W, K, mask)
    R, _ = orc.logl(W, K, mask, w.astype(np.float64), "poisson", sigma=0.4, phi=phi)
    Pm, lse, tot = orc.softmax(R)
    Mnum, Wt, T = orc.mstep(Pm, K, mask, q, coords, Ls, "trilinear", phi=phi)
    Mc, avg, _ = orc.normalize_support(orc.compress(Mnum, Wt), Wt)
    assert np.allclose(out["P"], Pm) and np.allclose(out["lse"], lse)
    assert np.allclose(out["model"], Mc) and np.isclose(out["total"], tot)

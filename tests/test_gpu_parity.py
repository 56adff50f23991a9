"""GPU parity: each stage and whole iterations of libemc against the FP64 oracle, through the
C ABI.  Tolerances are written per test; the discrete decisions (validity, voxel indices) are
bit-identical by construction and compared exactly."""
import numpy as np
import pytest

import emc_amd
from emc_amd import workloads as wl

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def rel_logl(R_gpu, R_orc):
    """max over (s, i) of |dR| / max_s |R[s, i]|: error relative to the pattern's logL magnitude
    (a single logit can cross zero, so |dR| / |R| per entry is ill-conditioned)."""
    scale = np.maximum(np.abs(R_orc).max(0, keepdims=True), 1.0)
    return float(np.max(np.abs(R_gpu.astype(np.float64) - R_orc) / scale))


def metrics(R_gpu, R_orc, P_gpu, P_orc, M_gpu, M_orc):
    out = {}
    if R_gpu is not None:
        out["rel_logl"] = rel_logl(R_gpu, R_orc)
    if P_gpu is not None:
        l1 = np.abs(P_gpu.astype(np.float64) - P_orc).sum(0)
        out["resp_l1_max"] = float(l1.max())
        out["resp_l1_mean"] = float(l1.mean())
    if M_gpu is not None:
        out["model_rel_l2"] = float(np.linalg.norm(M_gpu - M_orc) / np.linalg.norm(M_orc))
    return out


@pytest.fixture(scope="module")
def small():
    """32^3 model, 3,240 rotations, 1,000 patterns (the full-size small workload)."""
    d = wl.make("L32_S3240_I1000")
    d["m0"] = wl.normalized(d["model0"])
    return d


def engine(d, prec="fp32", S=None, I=None, **kw):
    L = d["L"]
    S = d["q"].shape[0] if S is None else S
    I = d["K"].shape[0] if I is None else I
    e = emc_amd.Engine(L, S, I, precision=prec, **kw)
    e.setup(d["q"][:S], d["w"][:S], d["coords"], d["mask"], d["K"][:I], d["m0"].astype(np.float32))
    return e


@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_expand_matches_oracle(orc, small, interp):
    log(f"[expand] interp={interp}")
    d = small
    with engine(d, "fp32", interp=interp) as e:
        W = e.expand(100, 200)
    Wo = orc.slices(d["m0"].astype(np.float32).astype(np.float64), d["L"], d["q"][100:300], d["coords"], interp)
    assert np.array_equal(W == -1, Wo == -1)            # identical validity decisions
    v = Wo != -1
    assert np.allclose(W[v], Wo[v], rtol=2e-6, atol=0)


@pytest.mark.parametrize("diff", ["true_poisson", "poisson", "absolute", "relative"])
@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_logl_fp32_matches_oracle(orc, small, diff, interp):
    log(f"[logl fp32] diff={diff} interp={interp}")
    d = small
    S, I = 700, 300
    sigma = 0.3
    with engine(d, "fp32", S=S, I=I, diff=diff, interp=interp) as e:
        e.logl(sigma)
        R = e.get_logl()
    M32 = d["m0"].astype(np.float32).astype(np.float64)
    W = orc.slices(M32, d["L"], d["q"][:S], d["coords"], interp)
    K = d["K"][:I].astype(np.float64)
    phi0 = orc.fluence(W, K, d["mask"]) if diff == "true_poisson" else 1.0
    Ro, _ = orc.logl(W, K, d["mask"], d["w"][:S].astype(np.float64), diff, sigma=sigma, phi0=phi0)
    err = rel_logl(R, Ro)
    log(f"   rel logL err {err:.2e}, max abs err {np.abs(R - Ro).max():.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("diff", ["poisson", "absolute", "relative"])
def test_scaled_logl_fp32_matches_oracle(orc, small, diff):
    """known_intensity = 0: per (rotation, pattern) scaling phi = sum K^2 / sum K W (emc_cuda.cu:438-455)
    divides the counts in the chi^2 logits (emc_cuda.cu:160-216)."""
    log(f"[logl scaled] diff={diff}")
    d = small
    S, I = 700, 300
    sigma = 0.3
    with engine(d, "fp32", S=S, I=I, diff=diff, scaling=1, slice_chunk=256) as e:
        e.logl(sigma)
        R = e.get_logl()
    M32 = d["m0"].astype(np.float32).astype(np.float64)
    W = orc.slices(M32, d["L"], d["q"][:S], d["coords"], "trilinear")
    K = d["K"][:I].astype(np.float64)
    phi = orc.scaling(W, K, d["mask"])
    Ro, _ = orc.logl(W, K, d["mask"], d["w"][:S].astype(np.float64), diff, sigma=sigma, phi=phi)
    R1, _ = orc.logl(W, K, d["mask"], d["w"][:S].astype(np.float64), diff, sigma=sigma)
    err = rel_logl(R, Ro)
    log(f"   rel logL err {err:.2e} (scaling moves the logits by {rel_logl(R1, Ro):.2e})")
    assert err < 1e-5
    assert rel_logl(R1, Ro) > 100 * err  # the scaled logits are not the unscaled ones


def test_scaling_rejected_where_unsupported(small):
    for kw in ({"precision": "bf16x2", "diff": "true_poisson"}, {"precision": "fp32", "diff": "true_poisson"}):
        with pytest.raises(emc_amd.EmcError):
            emc_amd.Engine(small["L"], 100, 50, scaling=1, **kw)


@pytest.mark.parametrize("scaling", [0, 1])
@pytest.mark.parametrize("prec", ["bf16x2", "bf16x3"])
@pytest.mark.parametrize("diff", ["poisson", "absolute", "relative"])
def test_chi2_logl_mfma_matches_oracle(orc, small, diff, prec, scaling):
    """The chi^2 kinds on the matrix cores: D expanded into sum a W^2 - 2 ip sum a W K + ip^2 sum a K^2
    (split-BF16 slice operands, exact BF16 counts and squared counts), with and without the
    intensity scaling's two contractions."""
    log(f"[logl chi2 mfma] diff={diff} prec={prec} scaling={scaling}")
    d = small
    S, I = 700, 300
    sigma = 0.3
    with engine(d, prec, S=S, I=I, diff=diff, scaling=scaling, slice_chunk=256) as e:
        e.logl(sigma)
        R = e.get_logl()
    M32 = d["m0"].astype(np.float32).astype(np.float64)
    W = orc.slices(M32, d["L"], d["q"][:S], d["coords"], "trilinear")
    K = d["K"][:I].astype(np.float64)
    phi = orc.scaling(W, K, d["mask"]) if scaling else None
    Ro, _ = orc.logl(W, K, d["mask"], d["w"][:S].astype(np.float64), diff, sigma=sigma, phi=phi)
    err = rel_logl(R, Ro)
    log(f"   rel logL err {err:.2e}, max abs err {np.abs(R - Ro).max():.2e}")
    assert err < (1e-4 if prec == "bf16x2" else 1e-5)


@pytest.mark.parametrize("prec", ["bf16x2", "bf16x3"])
def test_chi2_mfma_iterations_match_oracle(orc, small, prec):
    """Two iterations with the reference's own emc.conf choices (chi^2 "poisson", known_intensity =
    false) on the matrix cores; the M-step contracts the split P / phi."""
    log(f"[iterate chi2 mfma] prec={prec}")
    d = small
    L = d["L"]
    K = d["K"].astype(np.float64)
    w64 = d["w"].astype(np.float64)
    m = d["m0"].astype(np.float32)
    sig = 0.5
    with engine(d, prec, diff="poisson", scaling=1, sigma_start=sig, sigma_final=sig, blur_sigma=0.5) as e:
        for it in range(2):
            e.set_model(m)
            st = e.iterate(it)
            M, _ = e.get_model()
            P = e.get_resp()
            ref = orc.iterate(m.astype(np.float64), L, d["q"], w64, K, d["mask"], d["coords"], "trilinear",
                              "poisson", sigma=sig, blur_sigma=0.5, want_P=True, scaling=True)
            mt = metrics(None, None, P, ref["P"], M, ref["model"])
            rel_tot = abs(st["total_logl"] - ref["total"]) / abs(ref["total"])
            log(f"   it {it}: {mt} total logL rel {rel_tot:.2e}")
            assert rel_tot < 1e-5
            assert mt["resp_l1_max"] < 5e-3
            assert mt["model_rel_l2"] < 1e-4
            m = ref["model"].astype(np.float32)


@pytest.mark.parametrize("prec", ["bf16x2", "bf16x3"])
def test_logl_mfma_matches_oracle(orc, small, prec):
    log(f"[logl mfma] prec={prec}")
    d = small
    S, I = 1000, 400
    with engine(d, prec, S=S, I=I) as e:
        e.logl()
        R = e.get_logl()
    M32 = d["m0"].astype(np.float32).astype(np.float64)
    W = orc.slices(M32, d["L"], d["q"][:S], d["coords"], "trilinear")
    K = d["K"][:I].astype(np.float64)
    phi0 = orc.fluence(W, K, d["mask"])
    Ro, _ = orc.logl(W, K, d["mask"], d["w"][:S].astype(np.float64), "true_poisson", phi0=phi0)
    err = rel_logl(R, Ro)
    log(f"   rel logL err {err:.2e}, max abs err {np.abs(R - Ro).max():.2e} "
        f"(counts > 256 present: {d['K'][:I].max() > 256})")
    assert err < (1e-4 if prec == "bf16x2" else 1e-5)


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_responsibilities_match_oracle(orc, small, prec):
    log(f"[resp] prec={prec}")
    d = small
    S, I = 1200, 333  # ragged pattern count
    with engine(d, prec, S=S, I=I) as e:
        e.logl()
        tot = e.normalize_resp()
        P = e.get_resp()
        lse = e.get_lse()
        T = e.get_slice_weights()
    M32 = d["m0"].astype(np.float32).astype(np.float64)
    W = orc.slices(M32, d["L"], d["q"][:S], d["coords"], "trilinear")
    K = d["K"][:I].astype(np.float64)
    phi0 = orc.fluence(W, K, d["mask"])
    Ro, _ = orc.logl(W, K, d["mask"], d["w"][:S].astype(np.float64), "true_poisson", phi0=phi0)
    Po, lseo, toto = orc.softmax(Ro)
    assert np.allclose(P.sum(0), 1.0, atol=1e-4)
    l1 = np.abs(P - Po).sum(0)
    log(f"   resp L1 max {l1.max():.2e} mean {l1.mean():.2e}; lse max abs err {np.abs(lse - lseo).max():.2e}")
    assert l1.max() < 5e-3
    assert np.abs(lse - lseo).max() < 1e-5 * np.abs(lseo).max()
    assert abs(tot - toto) < 1e-5 * abs(toto)
    assert np.allclose(T, P.astype(np.float64).sum(1), rtol=1e-3, atol=1e-6)


@pytest.mark.parametrize("prec,insert", [("fp32", "atomic"), ("fp32", "binned"), ("bf16x2", "atomic"),
                                         ("bf16x2", "fused"), ("bf16x2", "binned")])
@pytest.mark.parametrize("interp", ["nearest", "trilinear"])
def test_maximize_and_compress_match_oracle(orc, small, prec, insert, interp):
    log(f"[mstep] prec={prec} insert={insert} interp={interp}")
    d = small
    S, I = 900, 250
    rng = np.random.default_rng(0)
    logits = rng.normal(scale=3.0, size=(S, I))
    Po = np.exp(logits - logits.max(0))
    Po /= Po.sum(0)
    Po[5] = 0.0                # a slice with T_s = 0 is skipped
    P32 = Po.astype(np.float32)
    with engine(d, prec, S=S, I=I, insert=insert, interp=interp, blur_sigma=0.5) as e:
        e.set_resp(P32)
        e.maximize()
        _, Wt = e.get_model()
        e.compress()
        M, _ = e.get_model()
    Pin = P32.astype(np.float64)
    Mnum, Wto, T = orc.mstep(Pin, d["K"][:I].astype(np.float64), d["mask"], d["q"][:S], d["coords"], d["L"], interp)
    Mo, _, _ = orc.normalize_support(orc.compress(Mnum, Wto), Wto)  # divide, normalise, blur
    Mo = orc.blur(Mo, d["L"], 0.5)
    assert np.array_equal(Wt > 0, Wto > 0)
    assert np.allclose(Wt, Wto, rtol=1e-4, atol=1e-6 * Wto.max())
    rel = np.linalg.norm(M - Mo) / np.linalg.norm(Mo)
    log(f"   Wt max rel {np.abs(Wt - Wto).max() / Wto.max():.2e}  model rel L2 {rel:.2e}")
    tol = 2e-5 if prec == "fp32" else 2e-3  # P enters BF16 hi+lo; the residual below 2^-16 is dropped
    assert rel < tol


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_full_iterations_match_oracle(orc, small, prec):
    """Three iterations of the small workload, each started from the oracle's previous model."""
    log(f"[iterate] prec={prec}")
    d = small
    L = d["L"]
    S, I = d["q"].shape[0], d["K"].shape[0]
    K = d["K"].astype(np.float64)
    w64 = d["w"].astype(np.float64)
    m = d["m0"].astype(np.float32)
    with engine(d, prec, blur_sigma=0.5) as e:
        for it in range(3):
            e.set_model(m)
            st = e.iterate(it)
            M, _ = e.get_model()
            P = e.get_resp()
            ref = orc.iterate(m.astype(np.float64), L, d["q"], w64, K, d["mask"], d["coords"], "trilinear",
                              "true_poisson", blur_sigma=0.5, want_P=True)
            mt = metrics(None, None, P, ref["P"], M, ref["model"])
            rel_tot = abs(st["total_logl"] - ref["total"]) / abs(ref["total"])
            log(f"   it {it}: {mt} total logL rel {rel_tot:.2e}")
            assert rel_tot < 1e-5
            assert mt["resp_l1_max"] < 5e-3
            assert mt["model_rel_l2"] < 1e-4
            m = ref["model"].astype(np.float32)


def test_sharp_posteriors_match_oracle(orc, small):
    """Once the model has formed, most (256-rotation band, pattern) blocks are negligible: the logL
    pass leaves their stores out and the softmax takes the workgroups with few patterns that are not
    negligible in its sparse form (one thread per such pattern, P staged in LDS).  An iteration from the
    model of the sixth, and the stage-wise responsibilities from that model (every logit stored, the
    same sparse softmax), against the oracle from the same model."""
    log("[sharp posteriors]")
    d = small
    L = d["L"]
    S, I = d["q"].shape[0], d["K"].shape[0]
    K = d["K"].astype(np.float64)
    w64 = d["w"].astype(np.float64)
    with engine(d, "bf16x2", blur_sigma=0.5, slice_chunk=1000) as e:
        for it in range(6):
            e.iterate(it)
        m = e.get_model()[0].copy()
        st = e.iterate(6)
        M, _ = e.get_model()
        P, T = e.get_resp(), e.get_slice_weights()
        e.set_model(m)
        e.logl()
        tot = e.normalize_resp()
        P2, lse2, T2 = e.get_resp(), e.get_lse(), e.get_slice_weights()
    ref = orc.iterate(m.astype(np.float64), L, d["q"], w64, K, d["mask"], d["coords"], "trilinear",
                      "true_poisson", blur_sigma=0.5, want_P=True)
    mt = metrics(None, None, P, ref["P"], M, ref["model"])
    l1 = np.abs(P2 - ref["P"]).sum(0)
    log(f"   iteration 6: {mt}, logit blocks left out {st['logits_skipped']:.3f}; stages: resp L1 max "
        f"{l1.max():.2e}, total logL rel {abs(tot - ref['total']) / abs(ref['total']):.2e}")
    assert st["logits_skipped"] > 0.3
    assert mt["resp_l1_max"] < 5e-3 and mt["model_rel_l2"] < 1e-4
    assert abs(st["total_logl"] - ref["total"]) < 1e-5 * abs(ref["total"])
    assert l1.max() < 5e-3 and abs(tot - ref["total"]) < 1e-5 * abs(ref["total"])
    assert np.array_equal(P, P2) and np.array_equal(T, T2)  # same model, same E-step arithmetic
    assert np.allclose(T, P.astype(np.float64).sum(1), rtol=1e-3, atol=1e-6)


def test_scaled_iterations_match_oracle(orc, small):
    """Two iterations with known_intensity = 0 and the reference's chi^2 ("poisson") diff: the
    scaling enters the logits and the M-step sums K P / phi (emc_atomic.cu:107-108)."""
    log("[iterate scaled]")
    d = small
    L = d["L"]
    S, I = d["q"].shape[0], d["K"].shape[0]
    K = d["K"].astype(np.float64)
    w64 = d["w"].astype(np.float64)
    m = d["m0"].astype(np.float32)
    sig = 0.5
    with engine(d, "fp32", diff="poisson", scaling=1, sigma_start=sig, sigma_final=sig, blur_sigma=0.5) as e:
        for it in range(2):
            e.set_model(m)
            st = e.iterate(it)
            M, _ = e.get_model()
            P = e.get_resp()
            ref = orc.iterate(m.astype(np.float64), L, d["q"], w64, K, d["mask"], d["coords"], "trilinear",
                              "poisson", sigma=sig, blur_sigma=0.5, want_P=True, scaling=True)
            mt = metrics(None, None, P, ref["P"], M, ref["model"])
            rel_tot = abs(st["total_logl"] - ref["total"]) / abs(ref["total"])
            log(f"   it {it}: {mt} total logL rel {rel_tot:.2e}")
            assert rel_tot < 1e-5
            assert mt["resp_l1_max"] < 5e-3
            assert mt["model_rel_l2"] < 1e-4
            m = ref["model"].astype(np.float32)


def test_sparse_photon_iterations_match_oracle(orc):
    """The sparse-count regime of the 128^3 / 300-photon workload (about 0.023 photons per active
    pixel, counts mostly 0 or 1) at the small shape: 18 photons over 780 pixels; byte counts on the
    MFMA path, two iterations against the oracle."""
    d = wl.make("L32_S3240_I1000", photons=18.0)
    K = d["K"]
    zeros = float((K[:, d["mask"] != 0] == 0).mean())
    log(f"[sparse] max count {K.max()}, zero fraction {zeros:.3f}")
    assert zeros > 0.95 and K.max() <= 16
    L = d["L"]
    m = wl.normalized(d["model0"]).astype(np.float32)
    K64 = K.astype(np.float64)
    w64 = d["w"].astype(np.float64)
    with emc_amd.Engine(L, d["q"].shape[0], K.shape[0], precision="bf16x2", blur_sigma=0.5) as e:
        e.setup(d["q"], d["w"], d["coords"], d["mask"], K, m)
        for it in range(2):
            e.set_model(m)
            st = e.iterate(it)
            M, _ = e.get_model()
            P = e.get_resp()
            ref = orc.iterate(m.astype(np.float64), L, d["q"], w64, K64, d["mask"], d["coords"], "trilinear",
                              "true_poisson", blur_sigma=0.5, want_P=True)
            mt = metrics(None, None, P, ref["P"], M, ref["model"])
            rel_tot = abs(st["total_logl"] - ref["total"]) / abs(ref["total"])
            log(f"   it {it}: {mt} total logL rel {rel_tot:.2e}")
            assert rel_tot < 1e-5
            assert mt["resp_l1_max"] < 5e-3
            assert mt["model_rel_l2"] < 1e-4
            m = ref["model"].astype(np.float32)


def test_chunked_rotations_equal_unchunked(small):
    log("[chunk]")
    d = small
    S, I = 1000, 200
    outs = []
    for chunk in (0, 333):
        with engine(d, "bf16x2", S=S, I=I, slice_chunk=chunk) as e:
            e.logl()
            R = e.get_logl()
            e.normalize_resp()
            e.maximize()
            e.compress()
            outs.append((R, e.get_model()[0]))
    assert np.allclose(outs[0][0], outs[1][0], rtol=1e-6, atol=1e-3)
    assert np.allclose(outs[0][1], outs[1][1], rtol=1e-4, atol=1e-6)


def test_edge_cases_empty_patterns_and_missing_voxels(orc, small):
    """All-zero patterns, a model with missing (-1) voxels, nearest interpolation."""
    log("[edge]")
    d = dict(small)
    S, I = 300, 40
    K = d["K"][:I].copy()
    K[3] = 0
    K[17] = 0
    m = d["m0"].astype(np.float32).copy()
    m.reshape(32, 32, 32)[:, :, :5] = -1.0
    d["K"] = K
    d["m0"] = m
    with engine(d, "fp32", S=S, I=I, interp="nearest") as e:
        e.logl()
        R = e.get_logl()
        e.normalize_resp()
        P = e.get_resp()
    W = orc.slices(m.astype(np.float64), 32, d["q"][:S], d["coords"], "nearest")
    Kd = K.astype(np.float64)
    phi0 = orc.fluence(W, Kd, d["mask"])
    Ro, _ = orc.logl(W, Kd, d["mask"], d["w"][:S].astype(np.float64), "true_poisson", phi0=phi0)
    Po, _, _ = orc.softmax(Ro)
    assert rel_logl(R, Ro) < 1e-5
    assert np.abs(P - Po).sum(0).max() < 1e-3
    # an empty pattern's posterior is the rotation prior reweighted by exp(-phi0 sum W_s)
    assert np.allclose(P[:, 3], Po[:, 3], atol=1e-5)


def test_invalid_inputs_are_rejected(small):
    d = small
    with pytest.raises(emc_amd.EmcError):
        emc_amd.Engine(32, 10, 10, precision="bf16x2", diff="true_poisson", scaling=1)
    with emc_amd.Engine(32, 60, 10) as e:
        with pytest.raises(emc_amd.EmcError):
            e.logl()  # nothing set
        bad = np.zeros(32 * 32, np.int32)
        bad[0] = 2
        with pytest.raises(emc_amd.EmcError):
            e.set_mask(bad)
        with pytest.raises(emc_amd.EmcError):
            e.set_rotations(np.zeros((60, 4), np.float32), np.ones(60, np.float32))  # not unit


def test_byte_counts_equal_bf16_counts(small):
    """Counts stored as bytes and widened in registers give bit-identical logits and models."""
    log("[count bytes]")
    d = dict(small)
    keep = d["K"].max(1) <= 255
    d["K"] = d["K"][keep][:300]
    outs = []
    for cb in (0, 2):
        with engine(d, "bf16x2", S=800, I=300, count_bytes=cb) as e:
            e.logl()
            R = e.get_logl()
            e.normalize_resp()
            e.maximize()
            e.compress()
            outs.append((R, e.get_model()[0]))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.allclose(outs[0][1], outs[1][1], rtol=1e-5, atol=1e-7)  # insertion order of float atomics varies


@pytest.mark.parametrize("chunk,diff", [(0, "true_poisson"), (300, "true_poisson"), (0, "poisson")])
def test_dma_staging_equals_register_staging(small, chunk, diff):
    """The MFMA contractions with operands staged in LDS by direct global->LDS loads give the same
    logits, log-sum-exps, responsibilities and (dense M-step) model, bit for bit, as the
    register-staged kernels (same products, same order): unchunked and over rotation chunks (band
    offsets of the softmax partials), and for the chi^2 kind with intensity scaling (its four
    contractions through the store epilogue)."""
    log(f"[mfma staging] chunk={chunk} diff={diff}")
    d = dict(small)
    d["K"] = d["K"][:700]
    kw = dict(diff=diff, scaling=1) if diff != "true_poisson" else {}
    outs = []
    for st in (0, 1):
        with engine(d, "bf16x2", S=1300, I=700, slice_chunk=chunk, mfma_staging=st, mstep="dense", **kw) as e:
            e.logl()
            R = e.get_logl()
            e.normalize_resp()
            P, lse = e.get_resp(), e.get_lse()
            e.maximize()
            e.compress()
            outs.append((R, P, lse, e.get_model()[0]))
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


def test_sparse_mstep_with_dense_rows_equals_dense(small):
    """Rows with thousands of non-zero responsibilities (9,000 patterns, 70 % non-zero): the sparse
    form's LDS list fills and is flushed several times within one scan step; it still gives the
    dense contraction's model."""
    log("[sparse M-step, dense rows]")
    d = dict(small)
    S, I = 300, 9000
    rng = np.random.default_rng(3)
    d["K"] = d["K"][rng.integers(0, d["K"].shape[0], I)]
    logits = rng.normal(scale=2.0, size=(S, I))
    Pm = np.exp(logits - logits.max(0))
    Pm[rng.random((S, I)) < 0.3] = 0.0
    Pm /= np.maximum(Pm.sum(0), 1e-30)
    P32 = Pm.astype(np.float32)
    out = {}
    for mode in ("dense", "sparse"):
        with engine(d, "bf16x2", S=S, I=I, mstep=mode) as e:
            e.set_resp(P32)
            e.maximize()
            e.compress()
            out[mode] = e.get_model()[0]
    rel = float(np.linalg.norm(out["sparse"] - out["dense"]) / np.linalg.norm(out["dense"]))
    log(f"   model rel L2 sparse vs dense {rel:.2e}")
    assert rel < 1e-6


@pytest.mark.parametrize("counts", ["bytes", "bf16", "bf16_split"])
def test_sparse_mstep_equals_dense(orc, small, counts):
    """The sparse M-step (only the non-zero responsibilities of each rotation) gives the dense
    contraction's model up to the order of the sum over patterns, for byte counts, BF16 counts and
    split BF16 counts (> 256); and the oracle's model."""
    log(f"[sparse M-step] counts={counts}")
    d = dict(small)
    K = d["K"][:400].copy()
    kw = {}
    if counts == "bf16":
        kw["count_bytes"] = 2
    elif counts == "bf16_split":
        K = K * 40  # counts above 256 need a second exact BF16 term
        assert K.max() > 256
    d["K"] = K
    engines = {mode: engine(d, "bf16x2", I=400, mstep=mode, blur_sigma=0.5, **kw) for mode in ("dense", "sparse", "auto")}
    try:
        m = d["m0"].astype(np.float32)
        for it in range(2):  # every form starts each iteration from the same model
            out = {}
            for mode, e in engines.items():
                e.set_model(m)
                st = e.iterate(it)
                out[mode] = e.get_model()[0], st
            for mode in ("sparse", "auto"):
                Ma, Mb = out[mode][0].astype(np.float64), out["dense"][0].astype(np.float64)
                rel = float(np.linalg.norm(Ma - Mb) / np.linalg.norm(Mb))
                dw = abs(out[mode][1]["weight_sum"] - out["dense"][1]["weight_sum"]) / out["dense"][1]["weight_sum"]
                log(f"   it {it} {mode}: model rel L2 vs dense {rel:.2e}, weight sum rel {dw:.2e}")
                assert rel < 1e-6 and dw < 1e-6
            m = out["dense"][0]
    finally:
        for e in engines.values():
            e.close()


@pytest.mark.parametrize("mstep", ["sparse", "auto"])
def test_reused_context_drops_the_lo_counts(small, mstep):
    """A context that once held counts above 256 (a BF16 hi + lo pair) and then gets counts <= 256
    with BF16 counts (count_bytes = 2) must not add the old lo terms anywhere: the sparse M-step
    and the per-pattern constants key on the current pattern set; same U / model and lse as a
    fresh context."""
    log(f"[reuse] lo counts dropped, M-step {mstep}")
    d = dict(small)
    S, I = 600, 200
    K = d["K"][:I]
    assert K.max() <= 256
    rng = np.random.default_rng(5)
    logits = rng.normal(scale=3.0, size=(S, I))
    Pm = np.exp(logits - logits.max(0))
    Pm[Pm < 1e-3] = 0.0  # sparse enough for the sparse form under AUTO
    P32 = (Pm / Pm.sum(0)).astype(np.float32)
    out = {}
    for reused in (False, True):
        with engine(dict(d, K=K), "bf16x2", S=S, I=I, count_bytes=2, mstep=mstep) as e:
            if reused:
                e.set_patterns(K * 40)   # max > 256: the lo planes are written
                e.logl()
                e.normalize_resp()
                e.set_patterns(K)        # back to one exact BF16 term
            e.logl()
            e.normalize_resp()
            lse = e.get_lse()
            e.set_resp(P32)
            e.maximize()
            e.compress()
            out[reused] = (lse, e.get_model()[0])
    assert np.array_equal(out[True][0], out[False][0])
    rel = float(np.linalg.norm(out[True][1] - out[False][1]) / np.linalg.norm(out[False][1]))
    log(f"   model rel L2 reused vs fresh {rel:.2e}")
    assert rel < 1e-6  # only the float atomics' order of the insertion may differ


def test_model_stats_between_logl_and_normalize(small):
    """emc_model_stats between emc_logl and emc_normalize_resp leaves the logL epilogue's softmax
    partials alone (the model reduction has its own scratch)."""
    log("[stage order] logl -> model_stats -> normalize_resp")
    d = small
    res = []
    for with_stats in (False, True):
        with engine(d, "bf16x2", S=1000, I=300) as e:
            e.logl()
            if with_stats:
                mx, avg = e.model_stats()
                assert mx > 0 and avg > 0
            tot = e.normalize_resp()
            res.append((tot, e.get_lse(), e.get_resp()))
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1]) and np.array_equal(res[0][2], res[1][2])


def _float_patterns(d, kind):
    """Non-integer pattern sets of the reference's own configuration: normalize_images output
    (each pattern scaled to sum P over the mask, emc_helper.cpp:389-418) and read_stride-binned
    means (pixel sum / mask sum, emc_helper.cpp:226-236: here the mean of a 2 x 2 block of two
    independent count images, so values are multiples of 1/4)."""
    K = d["K"].astype(np.float32)
    if kind == "normalized":
        return emc_amd.normalize_patterns(K, d["mask"])
    rng = np.random.default_rng(9)
    extra = rng.poisson(np.maximum(K, 0.0)).astype(np.float32)
    return ((K + extra) / 2.0 + rng.integers(0, 2, K.shape) * 0.25).astype(np.float32) * (d["mask"] != 0)


@pytest.mark.parametrize("kind", ["normalized", "binned_means"])
@pytest.mark.parametrize("diff,scaling", [("true_poisson", 0), ("poisson", 1)])
def test_float_patterns_on_mfma_match_oracle(orc, small, kind, diff, scaling):
    """Non-integer patterns run on the matrix cores as a BF16 hi + lo pair (no FP32 fallback):
    logits and a whole iteration against the FP64 oracle, for the true Poisson likelihood and for
    the reference's chi^2 ("poisson") with intensity scaling."""
    log(f"[float patterns] {kind}, diff={diff}, scaling={scaling}")
    d = dict(small)
    Kf = _float_patterns(d, kind)
    assert (Kf != np.floor(Kf)).any()
    S, I = 1200, 400
    Kf = Kf[:I]
    sig = 0.5
    kw = dict(diff=diff, scaling=scaling, sigma_start=sig, sigma_final=sig, blur_sigma=0.5)
    M32 = d["m0"].astype(np.float32)
    Kd = Kf.astype(np.float64)
    w64 = d["w"][:S].astype(np.float64)
    with engine(dict(d, K=Kf), "bf16x2", S=S, I=I, **kw) as e:
        e.logl(sig)
        R = e.get_logl()
        st = e.iterate(0)
        M, _ = e.get_model()
    W = orc.slices(M32.astype(np.float64), d["L"], d["q"][:S], d["coords"], "trilinear")
    phi = orc.scaling(W, Kd, d["mask"]) if scaling else None
    phi0 = orc.fluence(W, Kd, d["mask"]) if diff == "true_poisson" else 1.0
    Ro, _ = orc.logl(W, Kd, d["mask"], w64, diff, sigma=sig, phi=phi, phi0=phi0)
    ref = orc.iterate(M32.astype(np.float64), d["L"], d["q"][:S], w64, Kd, d["mask"], d["coords"], "trilinear",
                      diff, sigma=sig, blur_sigma=0.5, scaling=bool(scaling))
    err = rel_logl(R, Ro)
    rel = float(np.linalg.norm(M - ref["model"]) / np.linalg.norm(ref["model"]))
    dl = abs(st["total_logl"] - ref["total"]) / abs(ref["total"])
    log(f"   rel logL {err:.2e}, max abs {np.abs(R - Ro).max():.2e} nats, model rel L2 {rel:.2e}, "
        f"total logL rel {dl:.2e}")
    assert err < 1e-4
    assert rel < 1e-4 and dl < 1e-5


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_repeated_runs_are_bit_identical(small, prec):
    """Binned insertion sums in a fixed order (no atomics), the pixel column sums of the slicing
    statistics too, and every other reduction of the iteration has a fixed order: two runs of the
    same iterations give bit-identical models, responsibilities and log-likelihoods."""
    log(f"[determinism] {prec}, 3 iterations twice")
    d = small
    outs = []
    for run in range(2):
        with engine(d, prec, insert="binned", blur_sigma=0.5) as e:
            tots = [e.iterate(it)["total_logl"] for it in range(3)]
            M, Wt = e.get_model()
            outs.append((tots, M, Wt, e.get_lse(), e.get_resp()))
    a, b = outs
    assert a[0] == b[0]
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("chunk,skips,I", [(512, True, None), (1000, True, None), (700, False, None), (512, True, 333)])
def test_skipped_logit_stores_leave_the_iteration_unchanged(small, chunk, skips, I):
    """emc_iterate leaves out the stores of a 256-rotation band's logits for a pattern whose band
    maximum lies 128 nats below a maximum already found for it (keep_logits = 0, the default) and
    packs the kept ones of a tile into whole-line blocks; the softmax reads the blocks and takes the
    left-out bands as -inf.  The running maxima carry over the rotation chunks of a pass (at 3,240
    rotations one launch runs every tile at once, so chunks are what exercise the skipping here);
    chunks of a size that is not a multiple of 8 store every logit.  Over four iterations (the
    posteriors sharpen), models, weights, log-sum-exps, responsibilities, slice sums and total
    log-likelihoods equal those of keep_logits = 1 bit for bit, and the incomplete logits are not
    handed out.  A ragged pattern count (333) takes the responsibilities pass's element loads of the
    per-pattern state instead of its vector loads."""
    log(f"[logit store skipping] chunk={chunk} patterns={I or 'all'}")
    d = small
    outs, skipped = [], []
    for keep in (1, 0):
        with engine(d, "bf16x2", I=I, insert="binned", blur_sigma=0.5, slice_chunk=chunk, keep_logits=keep) as e:
            sts = [e.iterate(it) for it in range(4)]
            M, Wt = e.get_model()
            outs.append(([s["total_logl"] for s in sts], M, Wt, e.get_lse(), e.get_resp(), e.get_slice_weights()))
            skipped.append([s["logits_skipped"] for s in sts])
            if keep or not skips:
                e.get_logl()
            else:
                with pytest.raises(emc_amd.EmcError):
                    e.get_logl()
    log(f"   fraction of (band, pattern) logit blocks left out per iteration: {skipped[1]}")
    assert skipped[0] == [0.0] * 4
    assert (max(skipped[1]) > 0.1) if skips else skipped[1] == [0.0] * 4
    a, b = outs
    assert a[0] == b[0]
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


def test_compress_of_given_sums_matches_oracle(orc):
    """emc_set_mstep_sums + emc_compress: divide, normalise over the voxels with weight, blur, on
    M-step sums with empty (zero-weight) voxels and zero numerators, against the oracle."""
    log("[compress] given M-step sums")
    rng = np.random.default_rng(17)
    L = 24
    Wt = rng.uniform(0.0, 3.0, L ** 3).astype(np.float32)
    Wt[rng.random(L ** 3) < 0.2] = 0.0
    Mnum = (rng.uniform(0.0, 2.0, L ** 3) * Wt).astype(np.float32)
    Mnum[rng.random(L ** 3) < 0.1] = 0.0
    with emc_amd.Engine(L, 4, 4, blur_sigma=0.5) as e:
        e.set_mstep_sums(Mnum, Wt)
        avg = e.compress()
        M, W = e.get_model()
    Mo, avgo, _ = orc.normalize_support(orc.compress(Mnum.astype(np.float64), Wt.astype(np.float64)),
                                        Wt.astype(np.float64))
    Mo = orc.blur(Mo, L, 0.5)
    assert np.array_equal(W, Wt)
    assert abs(avg - avgo) <= 1e-6 * avgo
    miss = Wt == 0
    assert np.allclose(M[miss], -1.0 / avgo, rtol=1e-6) and np.allclose(Mo[miss], -1.0 / avgo)
    assert float(np.linalg.norm(M - Mo) / np.linalg.norm(Mo)) < 1e-6

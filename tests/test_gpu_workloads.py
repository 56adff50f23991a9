"""GPU checks at the benchmark workloads' full sizes, through the C ABI.

* 32^3 / 3,240 rotations / 1,000 patterns: ten chained iterations of the GPU path against ten
  chained iterations of the FP64 oracle from the same start (each side follows its own models).
* 64^3 / 25,680 / 20,000 (the plain FP32 and the split-BF16 forms): logits and responsibilities
  of the first 200 patterns over every rotation against the oracle (a pattern's logits depend on
  the others only through the global fluence, which the oracle takes from all patterns), and one
  whole iteration on the first 2,000 patterns.
* 128^3 / 50,100 / 50,000 at 300 photons: logits of 16 patterns over every rotation against the
  oracle (rotation-chunked on the host), and size-independent properties of a whole iteration
  (sum_s P = 1 for every pattern through sum_s T_s = I, finite log-sum-exps that add up to the
  total, rotation chunks of 5,000 equal to no chunking).
* 64^3 / 204,960 / 50,000: the grouped proposal iteration with every cell kept equals the full
  search.

Tolerances are written per test.  Logit errors are given two ways: relative to each pattern's
largest |logL| and as an absolute bound in nats (the softmax sees absolute differences).
"""
import numpy as np
import pytest

import emc_amd
from emc_amd import workloads as wl

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def rel_logl(R_gpu, R_orc):
    scale = np.maximum(np.abs(R_orc).max(0, keepdims=True), 1.0)
    return float(np.max(np.abs(R_gpu.astype(np.float64) - R_orc) / scale))


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


# ---- 32^3 / 3,240 / 1,000: ten chained iterations ---------------------------------------------

@pytest.fixture(scope="module")
def c0():
    d = wl.make("L32_S3240_I1000")
    d["m0"] = wl.normalized(d["model0"])
    return d


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_ten_chained_iterations_track_the_oracle(orc, c0, prec):
    d = c0
    L = d["L"]
    S, I = d["q"].shape[0], d["K"].shape[0]
    K = d["K"].astype(np.float64)
    w64 = d["w"].astype(np.float64)
    log(f"[chained] 32^3 / {S} / {I}, {prec}, 10 iterations")
    m_orc = d["m0"].copy()
    worst = {}
    with emc_amd.Engine(L, S, I, precision=prec, insert="atomic" if prec == "fp32" else "binned",
                        blur_sigma=0.5) as e:
        e.setup(d["q"], d["w"], d["coords"], d["mask"], d["K"], d["m0"].astype(np.float32))
        for it in range(10):
            st = e.iterate(it)
            M, _ = e.get_model()
            P = e.get_resp()
            ref = orc.iterate(m_orc, L, d["q"], w64, K, d["mask"], d["coords"], "trilinear", "true_poisson",
                              blur_sigma=0.5, want_P=True)
            m_orc = ref["model"]
            l1 = np.abs(P.astype(np.float64) - ref["P"]).sum(0)
            mt = {"model_rel_l2": rel_l2(M, m_orc), "resp_l1_max": float(l1.max()),
                  "resp_l1_mean": float(l1.mean()),
                  "total_logl_rel": abs(st["total_logl"] - ref["total"]) / abs(ref["total"]),
                  "fluence_rel": abs(st["fluence"] - ref["phi0"]) / ref["phi0"]}
            log(f"   it {it}: " + ", ".join(f"{k} {v:.2e}" for k, v in mt.items()))
            if it == 0:  # the same model in: phi0 (device sums vs orc_fluence) and the iteration agree
                assert mt["fluence_rel"] < 1e-6 and mt["total_logl_rel"] < 1e-5
                assert mt["model_rel_l2"] < 1e-4 and mt["resp_l1_max"] < 5e-3
            for k, v in mt.items():
                worst[k] = max(worst.get(k, 0.0), v)
    log(f"   worst over 10 iterations: {worst}")
    # Afterwards each side iterates on its own models.  From a random start some patterns' best
    # orientations nearly tie, and a difference at the rounding level moves such a pattern's whole
    # posterior (resp L1 up to 2), which then feeds back into the model: the two runs drift apart
    # without diverging.  Measured worst over 10 iterations: FP32 model 2.3e-3, total logL 1.5e-5,
    # mean resp L1 5.6e-3; BF16x2 5.8e-3, 2.1e-4, 5.4e-2.
    tol = {"fp32": (1e-2, 1e-4, 2e-2, 1e-3), "bf16x2": (2e-2, 1e-3, 0.15, 1e-3)}[prec]
    assert worst["model_rel_l2"] < tol[0]
    assert worst["total_logl_rel"] < tol[1]
    assert worst["resp_l1_mean"] < tol[2]
    assert worst["fluence_rel"] < tol[3]


# ---- 64^3 / 25,680 / 20,000 --------------------------------------------------------------------

@pytest.fixture(scope="module")
def c2():
    d = wl.make("L64_S25680_I20000")
    d["m0"] = wl.normalized(d["model0"]).astype(np.float32)
    return d


@pytest.fixture(scope="module")
def c2_oracle(orc, c2):
    """Oracle logits, responsibilities and lse of the first 200 patterns over every rotation, with
    the fluence of all 20,000 patterns."""
    d = c2
    log("[64^3 oracle] slices of 25,680 rotations, logits of 200 patterns")
    W = orc.slices(d["m0"].astype(np.float64), d["L"], d["q"], d["coords"], "trilinear")
    phi0 = orc.fluence(W, d["K"].astype(np.float64), d["mask"])
    R, _ = orc.logl(W, d["K"][:200].astype(np.float64), d["mask"], d["w"].astype(np.float64), "true_poisson",
                    phi0=phi0)
    del W
    P, lse, _ = orc.softmax(R.copy())
    return {"R": R, "P": P, "lse": lse, "phi0": phi0}


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec,tol_rel,tol_nats", [("fp32", 1e-5, 2e-2), ("bf16x2", 1e-4, 1e-2)])
def test_full_size_64_logits_of_200_patterns(c2, c2_oracle, prec, tol_rel, tol_nats):
    d, o = c2, c2_oracle
    S, I = d["q"].shape[0], d["K"].shape[0]
    log(f"[64^3 full size] {prec}: {S} rotations x {I} patterns, logits of the first 200")
    with emc_amd.Engine(d["L"], S, I, precision=prec, insert="atomic" if prec == "fp32" else "binned") as e:
        e.setup(d["q"], d["w"], d["coords"], d["mask"], d["K"], d["m0"])
        e.logl()
        R = e.get_logl_block(0, S, 0, 200)
        e.normalize_resp()
        P = e.get_resp_block(0, S, 0, 200)
        lse = e.get_lse()[:200]
    dR = np.abs(R.astype(np.float64) - o["R"])
    l1 = np.abs(P.astype(np.float64) - o["P"]).sum(0)
    mt = {"rel_logl": rel_logl(R, o["R"]), "max_abs_dlogl_nats": float(dR.max()),
          "resp_l1_max": float(l1.max()), "lse_max_abs": float(np.abs(lse - o["lse"]).max())}
    log("   " + ", ".join(f"{k} {v:.2e}" for k, v in mt.items()))
    assert mt["rel_logl"] < tol_rel
    assert mt["max_abs_dlogl_nats"] < tol_nats
    assert mt["lse_max_abs"] < tol_nats
    assert mt["resp_l1_max"] < 5e-3


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_full_size_64_iteration_on_2000_patterns(orc, c2, prec):
    d = c2
    S, I = d["q"].shape[0], 2000
    log(f"[64^3 iteration] {prec}: {S} rotations x {I} patterns, one iteration against the oracle")
    K = d["K"][:I]
    with emc_amd.Engine(d["L"], S, I, precision=prec, insert="atomic" if prec == "fp32" else "binned",
                        blur_sigma=0.5) as e:
        e.setup(d["q"], d["w"], d["coords"], d["mask"], K, d["m0"])
        st = e.iterate(0)
        M, Wt = e.get_model()
    ref = orc.iterate(d["m0"].astype(np.float64), d["L"], d["q"], d["w"].astype(np.float64), K.astype(np.float64),
                      d["mask"], d["coords"], "trilinear", "true_poisson", blur_sigma=0.5)
    mt = {"model_rel_l2": rel_l2(M, ref["model"]), "weight_rel_l2": rel_l2(Wt, ref["Wt"]),
          "total_logl_rel": abs(st["total_logl"] - ref["total"]) / abs(ref["total"]),
          "fluence_rel": abs(st["fluence"] - ref["phi0"]) / ref["phi0"]}
    log("   " + ", ".join(f"{k} {v:.2e}" for k, v in mt.items()))
    assert mt["fluence_rel"] < 1e-6
    assert mt["total_logl_rel"] < 1e-5
    assert mt["model_rel_l2"] < 1e-4
    assert mt["weight_rel_l2"] < 1e-4


@pytest.mark.timeout(600)
def test_full_size_64_formed_model_iteration_on_2000_patterns(orc, c2):
    """After six iterations the posteriors are sharp: the iteration leaves out most logit blocks and
    the softmax takes its sparse form in most workgroups.  The seventh iteration against the oracle's
    from the same model (responsibilities over every rotation, model, total logL)."""
    d = c2
    S, I = d["q"].shape[0], 2000
    log(f"[64^3 formed model] bf16x2: {S} rotations x {I} patterns, iteration 6 against the oracle")
    K = d["K"][:I]
    with emc_amd.Engine(d["L"], S, I, precision="bf16x2", blur_sigma=0.5) as e:
        e.setup(d["q"], d["w"], d["coords"], d["mask"], K, d["m0"])
        for it in range(6):
            e.iterate(it)
        m = e.get_model()[0].copy()
        st = e.iterate(6)
        M, Wt = e.get_model()
        P = e.get_resp()
    ref = orc.iterate(m.astype(np.float64), d["L"], d["q"], d["w"].astype(np.float64), K.astype(np.float64),
                      d["mask"], d["coords"], "trilinear", "true_poisson", blur_sigma=0.5, want_P=True)
    l1 = np.abs(P.astype(np.float64) - ref["P"]).sum(0)
    mt = {"model_rel_l2": rel_l2(M, ref["model"]), "weight_rel_l2": rel_l2(Wt, ref["Wt"]),
          "total_logl_rel": abs(st["total_logl"] - ref["total"]) / abs(ref["total"]),
          "resp_l1_max": float(l1.max()), "logit_blocks_left_out": st["logits_skipped"]}
    log("   " + ", ".join(f"{k} {v:.2e}" for k, v in mt.items()))
    assert mt["logit_blocks_left_out"] > 0.5
    assert mt["total_logl_rel"] < 1e-5
    assert mt["resp_l1_max"] < 5e-3
    assert mt["model_rel_l2"] < 1e-4
    assert mt["weight_rel_l2"] < 1e-4


# ---- 128^3 / 50,100 / 50,000, 300 photons ------------------------------------------------------

@pytest.fixture(scope="module")
def c3():
    d = wl.make("L128_S50100_I50000")
    d["m0"] = wl.normalized(d["model0"]).astype(np.float32)
    return d


@pytest.mark.timeout(900)
def test_full_size_128_logits_and_iteration_properties(orc, c3):
    d = c3
    L, S, I = d["L"], d["q"].shape[0], d["K"].shape[0]
    log(f"[128^3 full size] {S} rotations x {I} patterns, max count {d['K'].max()}")
    n_pat = 16
    # keep_logits: every logit of the iteration is stored (read below)
    with emc_amd.Engine(L, S, I, precision="bf16x2", blur_sigma=0.5, keep_logits=1) as e:
        e.setup(d["q"], d["w"], d["coords"], d["mask"], d["K"], d["m0"])
        st = e.iterate(0)
        R = e.get_logl_block(0, S, 0, n_pat)      # logits of the iteration's E-step (model m0)
        P = e.get_resp_block(0, S, 0, n_pat)
        lse = e.get_lse()
        T = e.get_slice_weights()
        M, Wt = e.get_model()
    log(f"   iteration: fluence {st['fluence']:.4f}, total logL {st['total_logl']:.6e}, "
        f"stage ms {st['stage_ms']['total']:.1f}")
    # properties of the whole iteration
    assert np.isfinite(lse).all() and np.isfinite(M).all()
    assert abs(lse.sum() - st["total_logl"]) <= 1e-9 * abs(st["total_logl"])
    assert abs(T.sum() - I) / I < 1e-5            # sum over rotations of every pattern's P is 1
    assert (T >= 0).all() and st["weight_sum"] > 0 and (M > 0).any()
    # the oracle's logits of the first n_pat patterns over every rotation, in rotation chunks,
    # with the device's fluence (the fluence itself is compared with orc_fluence at 32^3 and 64^3)
    Ro = np.zeros((S, n_pat))
    Ks = d["K"][:n_pat].astype(np.float64)
    m64 = d["m0"].astype(np.float64)
    for s0 in range(0, S, 5000):
        s1 = min(S, s0 + 5000)
        W = orc.slices(m64, L, d["q"][s0:s1], d["coords"], "trilinear")
        Ro[s0:s1], _ = orc.logl(W, Ks, d["mask"], d["w"][s0:s1].astype(np.float64), "true_poisson",
                                phi0=st["fluence"])
    Po, lseo, _ = orc.softmax(Ro.copy())
    dR = np.abs(R.astype(np.float64) - Ro)
    l1 = np.abs(P.astype(np.float64) - Po).sum(0)
    mt = {"rel_logl": rel_logl(R, Ro), "max_abs_dlogl_nats": float(dR.max()), "resp_l1_max": float(l1.max()),
          "lse_max_abs": float(np.abs(lse[:n_pat] - lseo).max())}
    log("   vs oracle: " + ", ".join(f"{k} {v:.2e}" for k, v in mt.items()))
    assert mt["rel_logl"] < 1e-4
    assert mt["max_abs_dlogl_nats"] < 1e-2 and mt["lse_max_abs"] < 1e-2
    assert mt["resp_l1_max"] < 5e-3
    # rotation chunks of 5,000 give the same iteration
    with emc_amd.Engine(L, S, I, slice_chunk=5000, precision="bf16x2", blur_sigma=0.5) as e:
        e.setup(d["q"], d["w"], d["coords"], d["mask"], d["K"], d["m0"])
        stc = e.iterate(0)
        Mc, _ = e.get_model()
    rc, rt = rel_l2(Mc, M.astype(np.float64)), abs(stc["total_logl"] - st["total_logl"]) / abs(st["total_logl"])
    log(f"   chunked (5,000) vs unchunked: model rel L2 {rc:.2e}, total logL rel {rt:.2e}")
    assert rc < 1e-5 and rt < 1e-7


# ---- 64^3 / 204,960 / 50,000 (learned proposal) ---------------------------------------------------

@pytest.mark.timeout(900)
def test_full_size_proposal_with_every_cell_equals_full_search():
    d = wl.make("L64_S204960_I50000")
    cq, _ = emc_amd.rotations(d["cfg"]["coarse_n"])
    L, S, I, C = d["L"], d["q"].shape[0], d["K"].shape[0], cq.shape[0]
    m0 = wl.normalized(d["model0"]).astype(np.float32)
    log(f"[204,960 rotations] {S} x {I}, {C} cells: grouped iteration with every cell kept vs full search")
    out = {}
    for mode in ("full", "grouped"):
        with emc_amd.Engine(L, S, I, precision="bf16x2", blur_sigma=0.5) as e:
            e.setup(d["q"], d["w"], d["coords"], d["mask"], d["K"], m0)
            if mode == "grouped":
                # every pattern's predicted mass is below min_mass: each is scored over every cell
                e.enable_proposal(cq, top_k=8, recall=0, warmup_full=1, full_every=1000, epochs=1,
                                  min_mass=2.0, max_fallback=1.0)
            s0 = e.iterate(0)
            s1 = e.iterate(1)
            out[mode] = (e.get_model()[0], s1, e.get_lse(), e.get_slice_weights())
            log(f"   {mode}: iteration 1 search {s1['search']}, work fraction {s1['work_fraction']:.3f}, "
                f"{s1['stage_ms']['total']:.1f} ms")
            assert s0["search"] == "full"
            assert s1["search"] == ("proposal" if mode == "grouped" else "full")
    (ma, sa, la, Ta), (mb, sb, lb, Tb) = out["grouped"], out["full"]
    assert sa["n_fallback"] == I and abs(sa["work_fraction"] - 1.0) < 1e-12
    rel = rel_l2(ma, mb.astype(np.float64))
    dl = abs(sa["total_logl"] - sb["total_logl"]) / abs(sb["total_logl"])
    dlse = la - lb
    log(f"   model rel L2 {rel:.2e}, total logL rel {dl:.2e} (grouped {sa['total_logl']:.9e}, full "
        f"{sb['total_logl']:.9e}); lse diff max {np.abs(dlse).max():.3e} mean {dlse.mean():.3e}; "
        f"sum lse {la.sum():.9e} / {lb.sum():.9e}; T rel {rel_l2(Ta, Tb):.2e}")
    # the centred logL leaves a total near zero (3.17 here), so the totals are compared in nats:
    # per pattern the two searches' lse agree to a few 1e-8
    assert rel < 1e-5 and np.abs(dlse).max() < 1e-6
    assert abs(sa["total_logl"] - sb["total_logl"]) < 1e-6 * I

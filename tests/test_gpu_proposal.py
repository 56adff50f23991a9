"""GPU: the learned orientation proposal through the C ABI -- network kernels (forward, gradients,
Adam) against the FP64 oracle, pattern features, the grouped iteration against the full search,
and what the trained network captures compared with the same network untrained."""
import numpy as np
import pytest

import emc_amd
from emc_amd import workloads as wl

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def bf16(a):
    """Round to the nearest BF16 (ties to even), as the device stores weights and features."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


@pytest.fixture(scope="module")
def d():
    """32^3 model, 3,240 fine rotations (n=4), 60 coarse cells (n=1), 1,000 patterns."""
    d = wl.make("L32_S3240_I1000")
    d["m_true"] = wl.normalized(d["model_true"]).astype(np.float32)
    d["cq"], _ = emc_amd.rotations(1)
    return d


def engine(d, I=None, model=None, **kw):
    I = d["K"].shape[0] if I is None else I
    e = emc_amd.Engine(d["L"], d["q"].shape[0], I, precision="bf16x2", **kw)
    e.setup(d["q"], d["w"], d["coords"], d["mask"], d["K"][:I], d["m_true"] if model is None else model)
    return e


def net_case(e, rng, B=50):
    """Random parameters with non-zero biases, features x [B][1024], soft targets t [B][C]."""
    W1, b1, W2, b2, W3, b3 = e.mlp_params()
    b1 = rng.normal(0, 0.1, b1.shape).astype(np.float32)
    b2 = rng.normal(0, 0.1, b2.shape).astype(np.float32)
    W3 = (W3 * 3.0).astype(np.float32)
    b3 = rng.normal(0, 0.3, b3.shape).astype(np.float32)
    e.set_mlp_params(W1, b1, W2, b2, W3, b3)
    x = rng.normal(size=(B, 1024)).astype(np.float32)
    t = rng.dirichlet(np.full(W3.shape[0], 0.3), size=B).astype(np.float32)
    params = dict(W1=bf16(W1), b1=b1, W2=bf16(W2), b2=b2, W3=bf16(W3), b3=b3)
    params = {k: v.astype(np.float64) for k, v in params.items()}
    return x, t, params


def test_mlp_forward_matches_oracle(orc, d):
    log("[mlp forward] H=128 C=60 B=50")
    rng = np.random.default_rng(5)
    with engine(d, I=256) as e:
        e.enable_proposal(d["cq"], hidden=128, batch=64)
        x, t, params = net_case(e, rng)
        p = e.mlp_forward(x)
    acts = orc.mlp_forward(bf16(x).astype(np.float64), params)
    err = float(np.abs(p - acts["p"]).max())
    log(f"  max |dp| = {err:.2e}, max p = {acts['p'].max():.3f}")
    assert np.allclose(p.sum(1), 1.0, atol=1e-5)
    assert err < 5e-3  # hidden activations are rounded to BF16 on the device


@pytest.mark.parametrize("rows,batch", [(50, 64), (200, 256)])
def test_mlp_gradients_match_oracle(orc, d, rows, batch):
    """The weight gradients reduce over the minibatch: a batch of 64 runs their products with a
    64-wide k-step, one of 256 with the 128-wide one."""
    log(f"[mlp gradients] H=128 C=60 B={rows} of a {batch}-row batch")
    rng = np.random.default_rng(6)
    with engine(d, I=256) as e:
        e.enable_proposal(d["cq"], hidden=128, batch=batch)
        x, t, params = net_case(e, rng, B=rows)
        loss, g = e.mlp_gradients(x, t)
    xb = bf16(x).astype(np.float64)
    acts = orc.mlp_forward(xb, params)
    go = orc.mlp_backward(xb, t, params, acts)
    lo = orc.mlp_loss(acts["p"], t)
    log(f"  loss gpu {loss:.6f} oracle {lo:.6f}")
    assert abs(loss - lo) / abs(lo) < 1e-2
    for name, a in zip(("W1", "b1", "W2", "b2", "W3", "b3"), g):
        b = go[name]
        rel = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        log(f"  grad {name}: rel L2 {rel:.2e}")
        assert rel < 3e-2, name  # deltas and activations are BF16 on the device


def test_adam_step_matches_oracle(orc, d):
    log("[adam] one step from zero moments")
    rng = np.random.default_rng(7)
    with engine(d, I=256) as e:
        e.enable_proposal(d["cq"], hidden=64, batch=64, lr=1e-3)
        x, t, _ = net_case(e, rng)
        p0 = e.mlp_params()
        _, g = e.mlp_gradients(x, t)
        e.mlp_train_step(x, t)
        p1 = e.mlp_params()
    for a0, ga, a1 in zip(p0, g, p1):
        w = a0.astype(np.float64).copy()
        m = np.zeros_like(w)
        v = np.zeros_like(w)
        orc.adam(w, ga.astype(np.float64), m, v, 1, lr=1e-3)
        assert float(np.abs(a1 - w).max()) < 1e-6


def test_features_are_pooled_log1p(d):
    log("[features] L=32 (no pooling)")
    with engine(d, I=64) as e:
        e.enable_proposal(d["cq"], hidden=64, batch=64)
        raw = e.proposal_features()
    K = d["K"][:64].reshape(64, 32, 32).astype(np.float32)
    K = K * d["mask"].reshape(32, 32)[None]
    ref = np.log1p(K).reshape(64, 1024)
    assert np.allclose(raw, ref, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("mstep", ["dense", "sparse", "auto"])
def test_grouped_iteration_with_every_cell_equals_full_search(d, mstep):
    """top_k = all 60 cells: the grouped path scores every (rotation, pattern) pair with the same
    operands and k order as the dense path, so the iteration reproduces the full search; with the
    dense and the sparse M-step forms."""
    log(f"[grouped == full] top_k = C = 60, M-step {mstep}")
    with engine(d, mstep=mstep) as a, engine(d, mstep=mstep) as b:
        a.enable_proposal(d["cq"], top_k=60, warmup_full=1, full_every=1000, epochs=1, hidden=64, batch=64)
        sa0, sb0 = a.iterate(0), b.iterate(0)
        assert sa0["search"] == "full"
        sa1, sb1 = a.iterate(1), b.iterate(1)
        log(f"  search {sa1['search']} work {sa1['work_fraction']:.3f} fallback {sa1['n_fallback']}")
        assert sa1["search"] == "proposal"
        assert sa1["n_fallback"] == 0 and abs(sa1["work_fraction"] - 1.0) < 1e-12
        ma, wa = a.get_model()
        mb, wb = b.get_model()
    rel = float(np.linalg.norm(ma.astype(np.float64) - mb) / np.linalg.norm(mb))
    dl = abs(sa1["total_logl"] - sb1["total_logl"]) / abs(sb1["total_logl"])
    log(f"  model rel L2 {rel:.2e}, total logL rel {dl:.2e}")
    assert rel < 1e-6
    assert dl < 1e-6  # the per-band FP32 exp of the online softmax differs with the band layout
    assert abs(sa1["weight_sum"] - sb1["weight_sum"]) / sb1["weight_sum"] < 1e-6


def test_trained_network_captures_mass_and_beats_untrained(d):
    """Captured responsibility mass of the top-8 of 60 cells: trained on the full-search targets
    vs the same network untrained (ablation), then the iteration policy end to end."""
    log("[proposal capture] top_k=8 of 60 cells, 1,000 patterns")
    with engine(d) as e, engine(d) as ref:
        e.enable_proposal(d["cq"], top_k=8, warmup_full=2, full_every=2, epochs=30, hidden=256, batch=128, seed=3)
        e.logl()
        e.normalize_resp()
        t = e.proposal_cell_targets()
        assert np.allclose(t.sum(1), 1.0, atol=1e-3)
        cells, mass = e.proposal_predict()
        cap_untrained = float(np.take_along_axis(t, cells, 1).sum(1).mean())
        loss = e.proposal_train(30)
        cells, mass = e.proposal_predict()
        cap_trained = float(np.take_along_axis(t, cells, 1).sum(1).mean())
        log(f"  captured: untrained {cap_untrained:.4f}, trained {cap_trained:.4f} "
            f"(predicted {float(mass.mean()):.4f}, loss {loss:.4f})")
        assert cap_trained > cap_untrained + 0.3
        assert cap_trained > 0.9
        # policy: iterations 0, 1 full (warm-up), 2 proposal, 3 full (measures the capture)
        for it in range(4):
            s = e.iterate(it)
            r = ref.iterate(it)
            ma, _ = e.get_model()
            mb, _ = ref.get_model()
            rel = float(np.linalg.norm(ma - mb) / np.linalg.norm(mb))
            log(f"  it {it}: {s['search']} work {s['work_fraction']:.3f} fallback {s['n_fallback']} "
                f"captured {s['captured_mass']:.4f} predicted {s['predicted_mass']:.4f} "
                f"loss {s['train_loss']:.4f} model rel L2 vs full {rel:.2e}")
            assert s["search"] == ("proposal" if it == 2 else "full")
            if it == 2:
                assert s["work_fraction"] < 0.5
            if it == 3:
                assert s["captured_mass"] > 0.9
            assert rel < 5e-2


def test_network_adds_capture_to_recall_at_lower_cost(d):
    """Ablation of the network inside the policy: the trained network's top-k plus the recall
    against the same policy with the network frozen at its random initialisation (Adam step
    1e-12, one pass), both from the random initial model, no fallback, so every proposal
    iteration searches exactly the union of the top-k and recalled cells.  The frozen arm is the
    recall plus k cells that carry no information about the pattern.  The trained arm captures
    more of the full search's responsibility mass while the model forms and searches fewer
    (rotation, pattern) pairs: its top-k overlaps the recalled cells."""
    k, R, n_it = 2, 2, 14
    log(f"[ablation] top_k={k} + recall={R} of 60 cells, trained vs frozen network, random start")
    m0 = wl.normalized(d["model0"]).astype(np.float32)
    arms = {}
    for arm, extra in (("trained", {}), ("frozen", dict(epochs=1, refresh_epochs=0, lr=1e-12))):
        with engine(d, model=m0, blur_sigma=0.5) as e:
            e.enable_proposal(d["cq"], top_k=k, recall=R, warmup_full=2, full_every=3, min_mass=0.0,
                              max_fallback=1.0, hidden=256, batch=128, seed=5, **extra)
            caps, work = [], []
            for it in range(n_it):
                s = e.iterate(it)
                if s["search"] == "full" and s["captured_mass"] >= 0:
                    caps.append(s["captured_mass"])
                elif s["search"] == "proposal":
                    work.append(s["work_fraction"])
        arms[arm] = (np.array(caps[:4]), float(np.mean(work)))
        log(f"  {arm}: captured at the full searches {np.round(caps, 4).tolist()}, "
            f"mean work fraction of the proposal iterations {np.mean(work):.4f}")
    (ct, wt), (cf, wf) = arms["trained"], arms["frozen"]
    assert ct.size == 4 and cf.size == 4
    assert ct.sum() > cf.sum() + 0.05
    assert wt < wf


def test_cell_targets_equal_the_cell_sums_of_the_responsibilities(d):
    """The network's targets Tt[i][c] = sum of P[s][i] over the fine rotations s of cell c, after
    full searches from the true model: the posteriors are sharp, so most 256-pattern row segments
    of P hold only zeros and the kernel does not read them."""
    log("[cell targets] 60 cells, 1,000 patterns, two full searches from the true model")
    cell = emc_amd.coarse_partition(d["q"], d["cq"])
    C = d["cq"].shape[0]
    with engine(d) as e:
        e.enable_proposal(d["cq"], top_k=8, warmup_full=10, full_every=2, epochs=1, hidden=64, batch=64)
        for it in range(2):
            assert e.iterate(it)["search"] == "full"
        t = e.proposal_cell_targets()
        P = e.get_resp().astype(np.float64)
    S, I = P.shape
    segs = np.pad(P, ((0, 0), (0, -I % 256))).reshape(S, -1, 256)
    zero = float((segs == 0).all(2).mean())
    want = np.zeros((I, C))
    for c in range(C):
        want[:, c] = P[cell == c].sum(0)
    err = float(np.abs(t - want).max())
    log(f"  zero row segments of P {zero:.3f}, max |Tt - cell sums of P| {err:.2e}")
    assert zero > 0.3
    assert np.allclose(t, want, rtol=1e-5, atol=1e-7)


def check_top_logit_cells(e, d, rec, R):
    """rec [I][R] holds, for each pattern, R distinct cells of largest max logit (the logits of
    the last full search; a tolerance of a few float roundings at the R-th value)."""
    cell = emc_amd.coarse_partition(d["q"], d["cq"])
    Lg = e.get_logl().astype(np.float64)
    C = d["cq"].shape[0]
    mx = np.full((C, Lg.shape[1]), -np.inf)
    for c in range(C):
        if (cell == c).any():
            mx[c] = Lg[cell == c].max(0)
    thr = -np.sort(-mx, axis=0)[R - 1]
    assert (rec >= 0).all()
    assert all(len(set(r)) == R for r in rec.tolist())
    got = np.take_along_axis(mx, rec.T.astype(np.int64), 0).min(0)
    assert (got >= thr - 1e-5 * np.maximum(1.0, np.abs(thr))).all()


def test_recall_follows_the_search(d):
    """Recalled cells: after a full search R cells of largest max logit; after a proposal
    iteration a subset of the cells it searched; the captured mass a full search measures counts
    the predicted and the recalled cells once each."""
    R, k = 3, 4
    log(f"[recall] top_k={k} recall={R} of 60 cells, 1,000 patterns")
    with engine(d) as e:
        e.enable_proposal(d["cq"], top_k=k, recall=R, warmup_full=1, full_every=3, epochs=2, hidden=64, batch=64,
                          min_mass=0.0, max_fallback=1.0, seed=4)
        assert (e.proposal_recall() == -1).all()
        assert e.iterate(0)["search"] == "full"
        rec = e.proposal_recall()
        check_top_logit_cells(e, d, rec, R)
        for it in (1, 2):  # proposal iterations (full_every=3 after one warm-up iteration)
            cells, _ = e.proposal_predict()
            s = e.iterate(it)
            assert s["search"] == "proposal"
            new = e.proposal_recall()
            for i in range(new.shape[0]):
                searched = set(cells[i].tolist()) | set(rec[i][rec[i] >= 0].tolist())
                got = new[i][new[i] >= 0].tolist()
                assert len(got) >= 1 and set(got) <= searched, (it, i, got, searched)
            log(f"  it {it}: work {s['work_fraction']:.3f}, recalled per pattern {float((new >= 0).sum(1).mean()):.2f}")
            rec = new
        cells, _ = e.proposal_predict()
        s = e.iterate(3)
        assert s["search"] == "full"
        t = e.proposal_cell_targets()
        cap = e.proposal_captured()
        want = np.zeros(t.shape[0])
        for i in range(t.shape[0]):
            cs = set(cells[i].tolist()) | set(rec[i][rec[i] >= 0].tolist())
            want[i] = sum(float(t[i, c]) for c in cs)
        log(f"  it 3: captured {s['captured_mass']:.4f} (top-{k} alone {float(np.take_along_axis(t, cells, 1).sum(1).mean()):.4f})")
        assert np.allclose(cap, want, atol=1e-5)
        assert abs(s["captured_mass"] - want.mean()) < 1e-6
        check_top_logit_cells(e, d, e.proposal_recall(), R)


@pytest.mark.parametrize("train_below", [0.0, 1.0])
def test_full_search_retrains_only_below_threshold(d, train_below):
    """After the warm-up, a full search retrains the network only when the capture it measured is
    below train_below: never with 0, always with 1; warm-up searches always train."""
    log(f"[retrain policy] train_below={train_below}")
    with engine(d) as e:
        e.enable_proposal(d["cq"], top_k=8, warmup_full=1, full_every=2, epochs=1, hidden=64, batch=64,
                          train_below=train_below, refresh_epochs=0, max_fallback=1.0)
        s0 = e.iterate(0)
        assert s0["search"] == "full" and s0["train_loss"] >= 0  # warm-up
        s1, s2 = e.iterate(1), e.iterate(2)
        assert s1["search"] == "proposal" and s2["search"] == "full" and s2["captured_mass"] >= 0
        log(f"  it 2 captured {s2['captured_mass']:.4f} train_loss {s2['train_loss']:.4f}")
        assert (s2["train_loss"] >= 0) == (train_below == 1.0)
        with pytest.raises(emc_amd.EmcError):
            e.enable_proposal(d["cq"], train_below=1.5)


def test_proposal_rejects_bad_use(d):
    log("[proposal errors]")
    with emc_amd.Engine(d["L"], d["q"].shape[0], 64, precision="fp32") as e:
        e.setup(d["q"], d["w"], d["coords"], d["mask"], d["K"][:64], d["m_true"])
        with pytest.raises(emc_amd.EmcError, match="UNSUPPORTED"):
            e.enable_proposal(d["cq"])
    with engine(d, I=64) as e:
        with pytest.raises(emc_amd.EmcError, match="INVALID"):
            e.enable_proposal(d["cq"], top_k=61)
        with pytest.raises(emc_amd.EmcError, match="INVALID"):
            e.enable_proposal(d["cq"], hidden=100)
        with pytest.raises(emc_amd.EmcError, match="INVALID"):
            e.enable_proposal(d["cq"], recall=9)
        with pytest.raises(emc_amd.EmcError, match="STATE"):
            e.proposal_predict()
        e.enable_proposal(d["cq"], hidden=64, batch=64)
        e.set_patterns(d["K"][:64])  # new patterns drop the proposal
        with pytest.raises(emc_amd.EmcError, match="STATE"):
            e.proposal_predict()

"""Host side of the learned proposal: the coarse-cell partition and the option defaults (no GPU)."""
import numpy as np
import pytest

import emc_amd


def brute_partition(fine, coarse):
    """argmax_c |fine_s . coarse_c| in FP64, lowest c on ties, in row blocks."""
    out = np.zeros(fine.shape[0], np.int32)
    f = fine.astype(np.float64)
    q = coarse.astype(np.float64)
    for b in range(0, f.shape[0], 4096):
        d = np.abs(f[b:b + 4096] @ q.T)
        out[b:b + 4096] = np.argmax(d, axis=1)
    return out


@pytest.mark.parametrize("nf,nc", [(4, 1), (8, 2)])
def test_partition_matches_brute_force(nf, nc):
    fq, _ = emc_amd.rotations(nf)
    cq, _ = emc_amd.rotations(nc)
    cell = emc_amd.coarse_partition(fq, cq)
    assert np.array_equal(cell, brute_partition(fq, cq))
    # every cell of the 600-cell-based coarse grid receives fine rotations of similar count
    counts = np.bincount(cell, minlength=cq.shape[0])
    assert counts.min() > 0
    assert counts.max() < 4 * counts.mean()


def test_partition_of_coarse_grid_is_identity():
    """Each coarse rotation is its own best match (|q.q| = 1, all others < 1)."""
    cq, _ = emc_amd.rotations(2)
    assert np.array_equal(emc_amd.coarse_partition(cq, cq), np.arange(cq.shape[0]))


def test_partition_sign_invariant():
    """q and -q are the same rotation: flipping signs does not change the cell."""
    fq, _ = emc_amd.rotations(4)
    cq, _ = emc_amd.rotations(1)
    a = emc_amd.coarse_partition(fq, cq)
    b = emc_amd.coarse_partition(-fq, cq)
    c = emc_amd.coarse_partition(fq, -cq)
    assert np.array_equal(a, b) and np.array_equal(a, c)


def test_partition_rejects_bad_input():
    with pytest.raises(emc_amd.EmcError):
        emc_amd.coarse_partition(np.zeros((0, 4), np.float32), np.zeros((1, 4), np.float32))


def test_proposal_defaults():
    o = emc_amd.proposal_options()
    assert (o.top_k, o.warmup_full, o.full_every, o.hidden, o.batch) == (8, 3, 5, 512, 1024)
    assert abs(o.lr - 1.5e-3) < 1e-9
    assert abs(o.min_mass - 0.99) < 1e-6 and o.synth_patterns == -1
    assert (o.recall, o.refresh_epochs, o.train_patterns) == (8, 1, 1)
    assert abs(o.train_below - 0.9999) < 1e-7
    with pytest.raises(KeyError):
        emc_amd.proposal_options(nonsense=1)

"""Pins the oracle's rotation sampling (rotations.c:77-458) without running the reference.

Checks: sample-count closed form, 600-cell combinatorics, the textbook vertex set, equal weights
at n=1, unit norms, weight sum, and invariance of (q, w) under left multiplication by elements
of the binary icosahedral group (the 600-cell's vertices), which any index or weight-formula slip
in the edge/face/cell subdivision breaks.
"""
import itertools

import numpy as np
import pytest

PHI = (1 + 5 ** 0.5) / 2


def textbook_600cell(order=(0.5, PHI / 2, 1 / (2 * PHI))):
    """The 120 unit quaternions: (+-1,0,0,0) perms, (+-1/2)^4, and even permutations of
    (+-a, +-b, +-c, 0) for (a, b, c) = order (Coxeter, Regular Polytopes, 8.7)."""
    v = []
    for i in range(4):
        for s in (-1, 1):
            e = [0.0] * 4
            e[i] = s
            v.append(e)
    for signs in itertools.product((-0.5, 0.5), repeat=4):
        v.append(list(signs))
    base = list(order) + [0.0]
    even = [p for p in itertools.permutations(range(4))
            if sum(1 for a in range(4) for b in range(a + 1, 4) if p[a] > p[b]) % 2 == 0]
    for p in even:
        for signs in itertools.product((-1, 1), repeat=3):
            e = [0.0] * 4
            vals = [base[0] * signs[0], base[1] * signs[1], base[2] * signs[2], 0.0]
            for k in range(4):
                e[p[k]] = vals[k]
            v.append(e)
    return np.array(v)


def set_close(a, b, tol=1e-6):
    d = np.abs(a[:, None, :] - b[None]).max(-1)
    return np.all(d.min(1) < tol) and np.all(d.min(0) < tol)


def upper(q):
    for x in q:
        if x > 0:
            return True
        if x != 0:
            return False
    return False


def qmul(a, b):
    a0, a1, a2, a3 = a
    b0, b1, b2, b3 = b.T
    return np.stack([a0 * b0 - a1 * b1 - a2 * b2 - a3 * b3,
                     a0 * b1 + a1 * b0 + a2 * b3 - a3 * b2,
                     a0 * b2 - a1 * b3 + a2 * b0 + a3 * b1,
                     a0 * b3 + a1 * b2 - a2 * b1 + a3 * b0], axis=1)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 10])
def test_count_and_combinatorics(orc, n):
    q, w, st = orc.rotations(n)
    assert q.shape[0] == 10 * (n + 5 * n ** 3) == orc.rotation_count(n)
    assert (st["edges"], st["faces"], st["cells"]) == (720, 1200, 600)
    assert st["kept"] == st["total"] // 2
    assert np.allclose(np.linalg.norm(q.astype(np.float64), axis=1), 1.0, atol=3e-7)
    assert np.all(w > 0)
    assert abs(w.astype(np.float64).sum() - 1.0) < 2e-4  # float running sum, as the reference
    assert all(upper(x) for x in q[:200])


def test_vertices_are_the_textbook_600cell(orc):
    q, w, _ = orc.rotations(1)
    tb = textbook_600cell()
    assert len(np.unique(np.round(tb, 9), axis=0)) == 120
    # the vertex set is the binary icosahedral group: closed under quaternion multiplication
    for g in tb[::7]:
        assert set_close(qmul(g, tb), tb)
    # every vertex has 12 nearest neighbours at 36 degrees (|vi + vj|^2 > 3, rotations.c:150)
    dots = tb @ tb.T
    assert np.all(((dots > np.cos(np.pi / 5) - 1e-9) & (dots < 1 - 1e-9)).sum(1) == 12)
    up = np.array([v for v in tb if upper(v)])
    assert up.shape == (60, 4) and q.shape == (60, 4)
    assert set_close(q.astype(np.float64), up)
    assert np.allclose(w, 1 / 60, rtol=1e-5)


@pytest.mark.parametrize("n", [2, 3, 4])
def test_weights_invariant_under_icosahedral_group(orc, n):
    q, w, _ = orc.rotations(n)
    q = q.astype(np.float64)
    full_q = np.concatenate([q, -q])
    full_w = np.concatenate([w, w]).astype(np.float64)
    group = textbook_600cell()
    rng = np.random.default_rng(n)
    for g in group[rng.choice(120, 6, replace=False)]:
        gq = qmul(g, q)
        # nearest neighbour of each image in the full (+-) list
        d2 = ((gq[:, None, :] - full_q[None, :, :]) ** 2).sum(-1) if q.shape[0] <= 1400 else None
        if d2 is None:
            continue
        j = d2.argmin(1)
        assert d2[np.arange(len(j)), j].max() < 1e-10
        assert np.allclose(full_w[j], w, rtol=2e-5)



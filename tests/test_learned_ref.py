"""CPU checks of the learned preconditioner's reference model (tests/learned_ref.py)."""
import numpy as np
import scipy.sparse as sp

import learned_ref as R


def tiny_problem(seed=0):
    """two triangles sharing an edge, one hinge; an SPD Hessian on that pattern"""
    tri = np.array([[0, 1, 2], [1, 3, 2]])
    hinge = np.array([[1, 2, 0, 3]])
    nV = 4
    row, col = R.elastic_pattern(nV, tri, hinge)
    rng = np.random.default_rng(seed)
    A = np.zeros((12, 12))
    for i, j in zip(row, col):
        if i < j:
            blk = rng.standard_normal((3, 3)) * 0.3
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = blk
            A[3 * j:3 * j + 3, 3 * i:3 * i + 3] = blk.T
    A += np.diag(rng.uniform(5, 10, 12))
    return sp.csr_matrix(A), tri, hinge, np.ones(nV), np.zeros(nV, bool)


def test_identity_factor_gives_zero_loss_in_expectation():
    """decoder at zero: G = I, so tr(G H^ G^T) = tr(H^) = n exactly and the loss is the
    Hutchinson error only; with many probes it goes to zero"""
    H, tri, hinge, m, dbc = tiny_problem()
    L, total = R.layout()
    w = np.zeros(total, np.float32)
    loss, g = R.loss_and_grad(H, m, dbc, tri, hinge, w, 4000, 5)
    assert abs(loss) < 0.03


def test_kaporin_loss_nonnegative_at_exact_trace():
    """Kaporin's K(M) = (tr M / n) / det(M)^(1/n) >= 1, so the exact-trace loss is >= 0 for
    any factor; check on random decoder outputs with a large probe count"""
    H, tri, hinge, m, dbc = tiny_problem(1)
    L, total = R.layout()
    rng = np.random.default_rng(3)
    w = (rng.standard_normal(total) * 0.05).astype(np.float32)
    loss, _ = R.loss_and_grad(H, m, dbc, tri, hinge, w, 4000, 9)
    assert loss > -0.03


def test_gradient_matches_central_differences():
    H, tri, hinge, m, dbc = tiny_problem(2)
    L, total = R.layout()
    rng = np.random.default_rng(4)
    # exactly representable in BF16 so that +-eps steps are not rounded away
    w = R.bf16((rng.standard_normal(total) * 0.05).astype(np.float32))
    loss, g = R.loss_and_grad(H, m, dbc, tri, hinge, w, 3, 21)
    o, n = L["bd"]
    for k in range(3):
        eps = 2.0 ** -6
        wp, wm = w.copy(), w.copy()
        wp[o + k] += eps
        wm[o + k] -= eps
        lp, _ = R.loss_and_grad(H, m, dbc, tri, hinge, wp, 3, 21)
        lm, _ = R.loss_and_grad(H, m, dbc, tri, hinge, wm, 3, 21)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[o + k]) <= 1e-3 * max(1.0, abs(fd)), k


def test_probe_vectors_are_rademacher_and_seeded():
    p = R.probes(7, 3, 999)
    assert set(np.unique(p)) == {-1.0, 1.0}
    assert abs(p.mean()) < 0.1
    assert np.array_equal(p, R.probes(7, 3, 999))
    assert not np.array_equal(p, R.probes(8, 3, 999))

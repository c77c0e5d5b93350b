"""Step-level invariants of the oracle (CPU only): Newton stopping rule, energy
decrease, intersection-free states, kappa adaptation bounds, velocity update."""
import numpy as np

import oracle as O
from cipc import scenes


def test_free_fall_matches_closed_form():
    """with no contact and the cloth at rest shape, one step is x + h v + h^2 g exactly
    up to the Newton tolerance (elastic energy is invariant under translation)"""
    sd = scenes.scene_c1()
    sd.params["PNTol"] = 1e-8 * sd.bbox_diag()
    orc = O.OracleScene(sd, threads=4)
    rc, st = orc.step()
    assert rc >= 3                                  # at least 3 Newton iterations
    x, v = orc.get_state()
    X = sd.X.ravel()
    h = sd.params["dt"]
    cloth = slice(0, 3 * sd.n_cloth)
    exp = X[cloth].reshape(-1, 3) + np.array([0, -9.81 * h * h, 0])
    assert np.abs(x[cloth].reshape(-1, 3) - exp).max() < 1e-7
    assert np.abs(v[cloth].reshape(-1, 3)[:, 1] + 9.81 * h).max() < 1e-5
    assert np.array_equal(x[3 * sd.n_cloth:], X[3 * sd.n_cloth:])   # Dirichlet nodes fixed


def test_contact_stack_steps():
    sd = scenes.scene_contact_stack()
    orc = O.OracleScene(sd, threads=4)
    k0 = orc.scalars()["kappa"]
    for _ in range(3):
        rc, st = orc.step()
        assert rc >= 3
        x, _ = orc.get_state()
        C = orc.constraint_set(x)
        if len(C):
            assert orc.dist2_list(x, C).min() > 0
        assert k0 <= st["kappa"] <= 100 * k0
        assert np.isfinite(x).all()

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


def test_static_solve_reaches_force_balance():
    """staticSolve (IMPLICIT_EULER.h:196-217, 646-666, 893-899): one step of a sheet pinned at
    its corners lands on the static equilibrium -- the elastic forces of the shell weighted
    by 1 (a dt = 1 copy of the scene gives that gradient) balance gravity at every free
    node -- independently of dt, and leaves the velocities untouched"""
    from test_oracle_strain import hammock
    out = []
    for dt in (0.01, 0.5):
        sd = hammock(0.0)
        sd.params["dt"] = dt
        sd.params["PNTol"] = 1e-8 * sd.bbox_diag()
        orc = O.OracleScene(sd)
        orc.set_static_solve(True)
        v0 = np.random.default_rng(1).uniform(-1, 1, 3 * sd.nV)
        x0, _ = orc.get_state()
        orc.set_state(x0, v0)
        rc, st = orc.step()
        assert rc >= 3
        x, v = orc.get_state()
        assert np.array_equal(v, v0)
        out.append(x)
        one = hammock(0.0)
        one.params["dt"] = 1.0
        ge, _ = O.OracleScene(one).gradient(x, np.zeros((0, 4), np.int32), 0.0)
        m = O.OracleScene(one).arrays()["mass"]
        g = np.array(sd.params["gravity"])
        free = np.ones(sd.nV, bool)
        free[sd.dbc_idx] = False
        resid = (ge.reshape(-1, 3) - m[:, None] * g)[free]
        assert np.abs(resid).max() <= 1e-6 * np.abs(m[:, None] * g).max()
        assert np.abs(x - x0).max() > 1e-3          # it sagged
    assert np.abs(out[0] - out[1]).max() <= 1e-9

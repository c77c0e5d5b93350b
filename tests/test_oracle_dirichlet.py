"""Pins for the oracle's moving Dirichlet nodes (FEM/BOUNDARY_CONDITION.h Step_Dirichlet,
IMPLICIT_EULER.h:304-401 and the penalty continuation :712-768), CPU only.

* Step_Dirichlet: targets rotated about the group's centre and axis by angVelDeg h, then
  translated by v h, against scipy's rotation (an independent implementation).
* unblocked motion: the nodes land exactly on their targets every step.
* blocked motion: a pusher triangle driven into a sheet lying 0.5 mm below it cannot move
  the whole way (the CCD cuts the Dirichlet displacement), the mass-weighted penalty pulls
  it on until 99% of the way is covered; the fixed corners stay exactly in place and every
  state is intersection free.
"""
import numpy as np
from scipy.spatial.transform import Rotation

import oracle as O
from cipc import scenes


def pinned_sheet(n=9, pusher_gap=None):
    """n x n sheet (0.3 m) with its 4 corners pinned; optionally a pusher triangle above
    its centre at `pusher_gap` (Dirichlet entries 4..6)."""
    X, T = scenes.cloth_grid(n, 0.0)
    X = 0.3 * X
    corners = [0, n - 1, n * (n - 1), n * n - 1]
    Xs, Ts = [X], [T]
    dbc = list(corners)
    if pusher_gap is not None:
        P = np.array([[-0.02, pusher_gap, -0.02], [0.02, pusher_gap, -0.015], [0.0, pusher_gap, 0.025]])
        Xs.append(P)
        Ts.append(np.array([[0, 1, 2]], dtype=np.int32) + len(X))
        dbc += [len(X), len(X) + 1, len(X) + 2]
    Xa = np.concatenate(Xs)
    Ta = np.concatenate(Ts).astype(np.int32)
    dbc = np.array(dbc, dtype=np.int32)
    sd = scenes.SceneData(Xa, Ta, dbc, Xa[dbc].copy(), dict(scenes.DEFAULT_PARAMS), "pinned_sheet")
    sd.params["PNTol"] = 1e-6 * sd.bbox_diag()
    sd.n_cloth = len(X)
    return sd


def test_step_dirichlet_closed_form():
    sd = pinned_sheet(pusher_gap=0.01)
    orc = O.OracleScene(sd)
    v, c, axis, w = np.array([0.3, -0.2, 0.1]), np.array([0.05, 0.0, -0.02]), np.array([0.0, 0.6, 0.8]), 45.0
    orc.dirichlet_add_motion(4, 7, v, c, axis, w)
    t0 = orc.dirichlet_targets()
    h = sd.params["dt"]
    R = Rotation.from_rotvec(axis * np.deg2rad(w) * h).as_matrix()
    exp = t0.copy()
    exp[4:7] = (t0[4:7] - c) @ R.T + c + v * h
    orc.step_dirichlet()
    assert np.abs(orc.dirichlet_targets() - exp).max() <= 1e-15
    # a second group: translation only
    orc.dirichlet_add_motion(0, 2, (0.0, 1.0, 0.0))
    t1 = orc.dirichlet_targets()
    orc.step_dirichlet()
    t2 = orc.dirichlet_targets()
    assert np.abs(t2[0:2] - (t1[0:2] + np.array([0.0, h, 0.0]))).max() <= 1e-15
    assert np.abs(t2[2:4] - t1[2:4]).max() == 0.0


def test_unblocked_motion_lands_on_targets():
    sd = pinned_sheet()
    orc = O.OracleScene(sd)
    orc.dirichlet_add_motion(0, 2, (0.0, 0.2, 0.0))      # lift two corners
    for _ in range(3):
        orc.step_dirichlet()
        rc, _ = orc.step()
        assert rc > 0
        x, _ = orc.get_state()
        assert np.array_equal(x.reshape(-1, 3)[sd.dbc_idx], orc.dirichlet_targets())


def test_blocked_motion_penalty_path():
    sd = pinned_sheet(pusher_gap=5e-4)
    orc = O.OracleScene(sd)
    orc.dirichlet_add_motion(4, 7, (0.0, -0.3, 0.0))     # 3 mm per step into a 0.5 mm gap
    for s in range(3):
        x0, _ = orc.get_state()
        orc.step_dirichlet()
        tgt = orc.dirichlet_targets()
        assert orc.ccd(x0, _dbc_disp(sd, x0, tgt), 1.0)[0] < 1.0    # the CCD blocks the motion
        rc, _ = orc.step()
        assert rc > 0, s
        x, _ = orc.get_state()
        X, X0 = x.reshape(-1, 3), x0.reshape(-1, 3)
        pin = sd.dbc_idx[:4]
        assert np.array_equal(X[pin], X0[pin])
        push = sd.dbc_idx[4:]
        gap = np.linalg.norm(tgt[4:] - X[push], axis=1)
        full = np.linalg.norm(tgt[4:] - X0[push], axis=1)
        assert np.sqrt((gap ** 2).sum() / (full ** 2).sum()) <= 0.01 + 1e-9
        C = orc.constraint_set(x)
        if len(C):
            assert orc.dist2_list(x, C).min() > 0
    # the sheet under the pusher was pushed down
    assert X[: sd.n_cloth, 1].min() < -1e-3


def _dbc_disp(sd, x, tgt):
    d = np.zeros_like(x).reshape(-1, 3)
    d[sd.dbc_idx] = tgt - x.reshape(-1, 3)[sd.dbc_idx]
    return d.ravel()


def strain_limited_translation(n=9, speed=0.3):
    """the pinned sheet with the 1.01 strain limit and all four corners translated along x
    at `speed` (3 mm per step at 0.3 m/s: moved alone, the corners would stretch their
    triangles by 8%, while the translated target shape itself is unstretched)"""
    sd = pinned_sheet(n)
    sd.params.update(kappa_s=(1e3, 0.0), s=(1.01, 0.0), sHat=(1.0, 1.0), collision=0)
    return sd


def test_motion_breaking_strain_limit_backtracks():
    """IMPLICIT_EULER.h:336-376: the Dirichlet displacement is halved until the moved nodes
    leave every triangle inside the strain limit; a cut motion takes the penalty path, which
    brings the corners to within 1% of their targets with every triangle feasible.  Moving
    the corners the whole way would start the step from an infeasible state."""
    sd = strain_limited_translation()
    orc = O.OracleScene(sd)
    orc.dirichlet_add_motion(0, 4, (0.3, 0.0, 0.0))
    probe = O.OracleScene(sd)
    for s in range(3):
        x0, _ = orc.get_state()
        orc.step_dirichlet()
        tgt = orc.dirichlet_targets()
        moved = x0.reshape(-1, 3).copy()
        moved[sd.dbc_idx] = tgt
        assert not probe.strain_energy_grad(moved.ravel())[0]    # moving only the corners is infeasible
        rc, _ = orc.step()
        assert rc > 0, s
        x, _ = orc.get_state()
        X, X0 = x.reshape(-1, 3), x0.reshape(-1, 3)
        gap = np.linalg.norm(tgt - X[sd.dbc_idx], axis=1)
        full = np.linalg.norm(tgt - X0[sd.dbc_idx], axis=1)
        assert np.sqrt((gap ** 2).sum() / (full ** 2).sum()) <= 0.01 + 1e-9
        assert probe.strain_energy_grad(x)[0]

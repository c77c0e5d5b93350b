"""Moving Dirichlet nodes on the GPU (through the C ABI) against the oracle.

Bars: Step_Dirichlet targets to 1e-15 of the oracle's; steps from the GPU's own state and
targets to bars.trajectory_bar; unblocked nodes exactly on their targets; blocked motion
(penalty continuation) within 1% of the prescribed displacement, fixed corners exactly
in place, every state intersection free.
"""
import numpy as np
import pytest

import oracle as O
from bars import trajectory_bar
from cipc import Simulator
from test_oracle_dirichlet import pinned_sheet, _dbc_disp, strain_limited_translation

pytestmark = pytest.mark.gpu


def test_step_dirichlet_matches_oracle(require_gpu):
    sd = pinned_sheet(pusher_gap=0.01)
    gpu = Simulator(sd)
    orc = O.OracleScene(sd)
    args = (4, 7, (0.3, -0.2, 0.1), (0.05, 0.0, -0.02), (0.0, 0.6, 0.8), 45.0)
    gpu.dirichlet_add_motion(*args)
    orc.dirichlet_add_motion(*args)
    gpu.dirichlet_add_motion(0, 2, (0.0, 1.0, 0.0))
    orc.dirichlet_add_motion(0, 2, (0.0, 1.0, 0.0))
    for _ in range(3):
        gpu.step_dirichlet()
        orc.step_dirichlet()
        assert np.abs(gpu.dirichlet_targets() - orc.dirichlet_targets()).max() <= 1e-15
    gpu.turn_dirichlet()
    t0 = gpu.dirichlet_targets()
    gpu.step_dirichlet()
    t1 = gpu.dirichlet_targets()
    assert np.abs(t1[0:2] - (t0[0:2] - np.array([0.0, sd.params["dt"], 0.0]))).max() <= 1e-15
    gpu.close()


def _run(sd, motion, steps, check):
    gpu = Simulator(sd, pcg_tol=1e-10)
    orc = O.OracleScene(sd)
    gpu.dirichlet_add_motion(*motion)
    for s in range(steps):
        x0, v0 = gpu.get_state()
        gpu.step_dirichlet()
        tgt = gpu.dirichlet_targets()
        orc.set_state(x0, v0)
        orc.set_kappa(gpu.scalars()["kappa"])
        orc.set_dirichlet_targets(tgt)
        st = gpu.step()
        rc, so = orc.step()
        assert rc > 0 and st["status"] == 0, (s, st, rc)
        xg, _ = gpu.get_state()
        xo, _ = orc.get_state()
        disp = np.abs(xo - x0).max()
        assert np.abs(xg - xo).max() <= trajectory_bar(sd, disp), (s, st["newton_iters"], so)
        check(gpu, orc, x0, xg, tgt)
        C = gpu.constraint_set(xg)
        if len(C):
            assert gpu.min_dist2(xg, C)[1] > 0
    gpu.close()
    return xg


def test_unblocked_motion(require_gpu):
    sd = pinned_sheet()

    def check(gpu, orc, x0, x, tgt):
        assert np.array_equal(x.reshape(-1, 3)[sd.dbc_idx], tgt)

    _run(sd, (0, 2, (0.0, 0.2, 0.0)), 3, check)


def test_blocked_motion_penalty(require_gpu):
    sd = pinned_sheet(pusher_gap=5e-4)

    def check(gpu, orc, x0, x, tgt):
        X, X0 = x.reshape(-1, 3), x0.reshape(-1, 3)
        assert gpu.ccd_stepsize(x0, _dbc_disp(sd, x0, tgt), 1.0)[0] < 1.0   # the penalty path ran
        assert np.array_equal(X[sd.dbc_idx[:4]], X0[sd.dbc_idx[:4]])
        push = sd.dbc_idx[4:]
        gap = np.linalg.norm(tgt[4:] - X[push], axis=1)
        full = np.linalg.norm(tgt[4:] - X0[push], axis=1)
        assert np.sqrt((gap ** 2).sum() / (full ** 2).sum()) <= 0.01 + 1e-9

    x = _run(sd, (4, 7, (0.0, -0.3, 0.0)), 3, check)
    assert x.reshape(-1, 3)[: sd.n_cloth, 1].min() < -1e-3


def test_motion_breaking_strain_limit(require_gpu):
    """the corners translated 3 mm per step with the 1.01 strain limit on: the Dirichlet
    displacement is halved until the moved corners leave every triangle feasible
    (IMPLICIT_EULER.h:336-376), the penalty then brings them within 1% of their targets"""
    sd = strain_limited_translation()
    probe = O.OracleScene(sd)

    def check(gpu, orc, x0, x, tgt):
        X, X0 = x.reshape(-1, 3), x0.reshape(-1, 3)
        moved = X0.copy()
        moved[sd.dbc_idx] = tgt
        assert not probe.strain_energy_grad(moved.ravel())[0]
        gap = np.linalg.norm(tgt - X[sd.dbc_idx], axis=1)
        full = np.linalg.norm(tgt - X0[sd.dbc_idx], axis=1)
        assert np.sqrt((gap ** 2).sum() / (full ** 2).sum()) <= 0.01 + 1e-9
        assert probe.strain_energy_grad(x)[0]

    _run(sd, (0, 4, (0.3, 0.0, 0.0)), 3, check)


def test_write_obj(require_gpu, tmp_path):
    """Write_TriMesh_Obj of the current state: %le vertices, 1-based faces"""
    sd = pinned_sheet(pusher_gap=0.01)
    gpu = Simulator(sd)
    gpu.step()
    x, _ = gpu.get_state()
    path = tmp_path / "frame.obj"
    gpu.write_obj(path)
    V, F = [], []
    for line in path.read_text().splitlines():
        tag, *vals = line.split()
        (V if tag == "v" else F).append([float(v) for v in vals] if tag == "v" else [int(v) for v in vals])
    assert np.allclose(np.array(V), x.reshape(-1, 3), rtol=1e-6, atol=1e-12)
    assert np.array_equal(np.array(F) - 1, sd.tris)
    gpu.close()

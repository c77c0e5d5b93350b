"""Boundary observables of the step on the GPU (through the C ABI) against the oracle:
staticSolve (IMPLICIT_EULER.h:196-217, 646-666, 893-899) -- the pinned sheet's static
equilibrium to bars.trajectory_bar, velocities untouched, friction refused -- and the
per-step text logs written under the output folder."""
import numpy as np
import pytest

import oracle as O
from bars import trajectory_bar
from cipc import Simulator
from cipc.sim import CipcError
from test_oracle_strain import hammock

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_static_solve_matches_oracle(require_gpu, precision):
    sd = hammock(0.0)
    sd.params["PNTol"] = 1e-8 * sd.bbox_diag()
    gpu = Simulator(sd, pcg_tol=1e-12, precision=precision)
    gpu.set_static_solve(True)
    orc = O.OracleScene(sd)
    orc.set_static_solve(True)
    x0, _ = gpu.get_state()
    v0 = np.random.default_rng(1).uniform(-1, 1, x0.size)
    gpu.set_state(x0, v0)
    orc.set_state(x0, v0)
    st = gpu.step()
    rc, so = orc.step()
    assert st["status"] == 0 and rc > 0
    xg, vg = gpu.get_state()
    xo, _ = orc.get_state()
    assert np.array_equal(vg, v0)
    assert np.abs(xg - xo).max() <= trajectory_bar(sd, np.abs(xo - x0).max()), (st["newton_iters"], so["newton_iters"])
    gpu.close()


def test_static_solve_refuses_friction(require_gpu):
    sd = hammock(0.0)
    sd.params.update(mu=0.3, epsv2=1e-6, fricIterAmt=1, collision=1)
    gpu = Simulator(sd)
    gpu.set_static_solve(True)
    with pytest.raises(CipcError) as e:
        gpu.step()
    assert e.value.code == -5
    gpu.close()


def test_step_text_logs(require_gpu, tmp_path):
    """residual.txt: one line per Newton iteration, the last L2Norm the step's final
    residual; counter.txt: "PNIter contacts" per step; stretch.txt: the largest singular
    value of the triangles' in-plane deformation (numpy SVD of the same F) and the mean of
    those above 1, the smallest and the mean of those below 1"""
    from cipc import scenes
    sd = scenes.scene_contact_stack(n=8)
    gpu = Simulator(sd)
    gpu.set_output_folder(tmp_path)
    stats = [gpu.step() for _ in range(2)]
    res = np.loadtxt(tmp_path / "residual.txt", ndmin=2)
    assert len(res) == sum(s["newton_iters"] for s in stats)
    assert res[-1, 0] == stats[-1]["newton_iters"]
    # the files hold %le (7 significant digits), as the reference writes them
    assert abs(res[-1, 4] - stats[-1]["last_residual"]) <= 1e-6 * abs(stats[-1]["last_residual"])
    assert abs(res[-1, 3] - stats[-1]["energy"]) <= 1e-6 * abs(stats[-1]["energy"])
    cnt = np.loadtxt(tmp_path / "counter.txt", ndmin=2)
    assert cnt.tolist() == [[s["newton_iters"], s["contacts"]] for s in stats]
    st = np.loadtxt(tmp_path / "stretch.txt", ndmin=2)
    x, _ = gpu.get_state()
    X = x.reshape(-1, 3)
    smax, smin = [], []
    dbc = np.zeros(sd.nV, bool)
    dbc[sd.dbc_idx] = True
    for t in sd.tris:
        if dbc[t].all():
            continue
        e1, e2 = X[t[1]] - X[t[0]], X[t[2]] - X[t[0]]
        l = np.linalg.norm(e1)
        A = np.array([[l, e1 @ e2 / l], [0.0, np.linalg.norm(np.cross(e1, e2)) / l]])
        r1, r2 = sd.X[t[1]] - sd.X[t[0]], sd.X[t[2]] - sd.X[t[0]]
        I = np.array([[r1 @ r1, r1 @ r2], [r1 @ r2, r2 @ r2]])
        B = np.linalg.cholesky(I).T                 # upper triangular, B^T B = rest metric
        s = np.linalg.svd(A @ np.linalg.inv(B), compute_uv=False)
        smax.append(s.max())
        smin.append(s.min())
    smax, smin = np.array(smax), np.array(smin)
    exp = [max(1.0, smax.max()), smax[smax > 1].mean() if (smax > 1).any() else 0.0,
           min(1.0, smin.min()), smin[smin < 1].mean() if (smin < 1).any() else 0.0]
    assert np.allclose(st[-1], exp, rtol=2e-6, atol=0)
    gpu.close()

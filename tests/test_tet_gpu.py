"""Volumetric tets inside the shell step on the GPU (through the C ABI) against the oracle.

Scene: a soft Neo-Hookean tet cube (384 tets) on the ground with a cloth sheet 0.8 mm
above its top.  Bars: lumped mass and contact set (the tets' surface triangles ahead of
the shell's) bit-exact; elastic energy 1e-10 relative, gradient 1e-10, Hessian with the
projected tet blocks 1e-9 (Jacobi vs QL eigensolvers in the singular basis); the
inversion-free step size 1e-12 relative; four steps from the GPU's own state to
bars.trajectory_bar, with no tet inverted and every state intersection free.
"""
import numpy as np
import pytest

import oracle as O
from bars import trajectory_bar
from cipc import Simulator, scenes

pytestmark = pytest.mark.gpu
NONE = np.zeros((0, 4), np.int32)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def tet_dets(sd, x):
    X = x.reshape(-1, 3)[sd.tets]
    return np.linalg.det(np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], -1))


def test_tet_terms_match_oracle(require_gpu):
    sd = scenes.scene_tet_cloth()
    gpu = Simulator(sd)
    orc = O.OracleScene(sd, threads=8)
    nTet, nBT = gpu.tet_sizes()
    assert nTet == len(sd.tets) and nBT == orc.nBT == len(orc.tet_surface()) + sd.nT
    assert np.array_equal(gpu.setup_arrays()["mass"], orc.arrays()["mass"])
    x, _ = gpu.get_state()
    C = gpu.constraint_set(x)
    assert len(C) > 100 and np.array_equal(C, orc.constraint_set(x))
    # a sheared, squashed state of the cube
    rng = np.random.default_rng(4)
    y = x.reshape(-1, 3).copy()
    cube = np.unique(sd.tets)
    y[cube, 0] += 0.05 * y[cube, 1]
    y[cube, 1] *= 0.9
    y[cube] += rng.uniform(-2e-3, 2e-3, (len(cube), 3))
    y = y.ravel()
    assert (tet_dets(sd, y) > 0).all()
    eg = gpu.elastic_energy(y)
    eo = orc.energy(y, None, NONE, 1.0)[0]
    assert abs(eg[0] - eo[0]) <= 1e-10 * abs(eo[0])
    go, _ = orc.gradient(y, NONE, 1.0)
    assert rel(gpu.elastic_gradient(y), go) <= 1e-10
    Hg = gpu.hessian(y, NONE, 1.0, which=1).toarray()
    Ho = orc.hessian(y, NONE, 1.0, which=1).toarray()
    assert rel(Hg, Ho) <= 1e-9
    # inversion-free step size along a compressing direction
    p = np.zeros_like(y).reshape(-1, 3)
    p[cube, 1] = -0.5 * (y.reshape(-1, 3)[cube, 1] - y.reshape(-1, 3)[cube, 1].min()) + rng.uniform(-1e-3, 1e-3, len(cube))
    p = p.ravel()
    ao = orc.inversion_free_stepsize(y, p, 10.0)
    ag = gpu.inversion_free_stepsize(y, p, 10.0)
    assert ao < 10.0
    assert abs(ag - ao) <= 1e-12 * ao
    # an inverted state is an invalid potential on both sides
    z = y.reshape(-1, 3).copy()
    z[cube, 1] *= -1
    assert np.isinf(gpu.elastic_energy(z.ravel())[0])
    gpu.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tet_cloth_trajectory(require_gpu, precision):
    sd = scenes.scene_tet_cloth()
    gpu = Simulator(sd, pcg_tol=1e-10, precision=precision)
    orc = O.OracleScene(sd, threads=8)
    for s in range(4):
        x0, v0 = gpu.get_state()
        orc.set_state(x0, v0)
        orc.set_kappa(gpu.scalars()["kappa"])
        st = gpu.step()
        rc, so = orc.step()
        assert rc > 0 and st["status"] == 0, (s, st, rc)
        xg, _ = gpu.get_state()
        xo, _ = orc.get_state()
        disp = np.abs(xo - x0).max()
        assert np.abs(xg - xo).max() <= trajectory_bar(sd, disp), (s, st["newton_iters"], so["newton_iters"])
        assert (tet_dets(sd, xg) > 0).all()
        C = gpu.constraint_set(xg)
        assert len(C) and gpu.min_dist2(xg, C)[1] > 0
    gpu.close()

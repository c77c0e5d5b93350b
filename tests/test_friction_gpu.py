"""Lagged friction on the GPU (through the C ABI) against the oracle.

Bars: friction stencils bit-exact; closest points, tangent bases and normal forces to
1e-12 relative (same formulas, contraction allowed outside the basis kernel); energy
1e-11 and gradient 1e-11 relative (summation order); Hessian 1e-10 relative on the free
rows/columns; steps (each from the GPU's own state) to the trajectory bar of
bars.trajectory_bar; the sliding sheet's momentum balance to 1e-4.
"""
import numpy as np
import pytest

import oracle as O
from bars import trajectory_bar
from cipc import Simulator, scenes
from test_oracle_friction import fric_scene

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)


@pytest.fixture(scope="module")
def stack(require_gpu):
    sd = scenes.scene_contact_stack()
    sd.params.update(mu=0.2, epsv2=1e-6)
    gpu = Simulator(sd)
    orc = O.OracleScene(sd, threads=4)
    x, _ = orc.get_state()
    C = orc.constraint_set(x)
    kappa = orc.scalars()["kappa"]
    return sd, gpu, orc, x, C, kappa


def test_friction_basis_matches_oracle(stack):
    sd, gpu, orc, x, C, kappa = stack
    FCg, FIg = gpu.friction_basis(x, C, kappa)
    FCo, FIo = orc.friction_basis(x, C, kappa)
    assert len(FCo) > 100
    assert np.array_equal(FCg, FCo)
    for k in range(9):
        assert rel(FIg[:, k], FIo[:, k]) <= 1e-12, k


@pytest.mark.parametrize("scale", [0.2, 3.0, 50.0])
def test_friction_energy_gradient_hessian(stack, scale):
    """at x = xn + noise of `scale` eps_v h: static, mixed and sliding stencils"""
    sd, gpu, orc, x, C, kappa = stack
    FCg, FIg = gpu.friction_basis(x, C, kappa)
    FCo, FIo = orc.friction_basis(x, C, kappa)
    ev = np.sqrt(1e-6) * sd.params["dt"]
    rng = np.random.default_rng(int(10 * scale))
    xx = x + rng.standard_normal(x.size) * ev * scale
    eg = gpu.friction_energy(xx, x)
    eo = orc.friction_energy(xx, x, FCo, FIo)
    assert abs(eg - eo) <= 1e-11 * abs(eo)
    assert rel(gpu.friction_gradient(xx, x), orc.friction_gradient(xx, x, FCo, FIo)) <= 1e-11
    Hg = gpu.friction_hessian(xx, x).toarray()
    Ho = orc.friction_hessian(xx, x, FCo, FIo).toarray()
    free = np.ones(sd.nV, bool)
    free[sd.dbc_idx] = False
    f3 = np.repeat(free, 3)
    Ho[~f3, :] = 0
    Ho[:, ~f3] = 0
    assert rel(Hg, Ho) <= 1e-10


def _trajectory(sd, steps, v0=None, precision="fp64"):
    gpu = Simulator(sd, pcg_tol=1e-10, precision=precision)
    orc = O.OracleScene(sd, threads=4)
    if v0 is not None:
        x, _ = gpu.get_state()
        gpu.set_state(x, v0.ravel())
    for s in range(steps):
        x0, v0_ = gpu.get_state()
        orc.set_state(x0, v0_)
        orc.set_kappa(gpu.scalars()["kappa"])
        st = gpu.step()
        rc, so = orc.step()
        assert rc > 0 and st["status"] == 0, (s, st, rc)
        xg, _ = gpu.get_state()
        xo, _ = orc.get_state()
        disp = np.abs(xo - x0).max()
        assert np.abs(xg - xo).max() <= trajectory_bar(sd, disp), (s, st["newton_iters"], so)
        C = gpu.constraint_set(xg)
        if len(C):
            assert gpu.min_dist2(xg, C)[1] > 0
    return gpu, st


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_friction_step_trajectory_stack(require_gpu, precision):
    """three-sheet contact stack with mu = 0.2: four steps, each from the GPU's state"""
    sd = scenes.scene_contact_stack()
    sd.params.update(mu=0.2, epsv2=1e-6)
    sd.params["PNTol"] = 1e-6 * sd.bbox_diag()
    _trajectory(sd, 4, precision=precision)


@pytest.mark.parametrize("fric_iter", [1, 0])
def test_friction_step_trajectory_sliding(require_gpu, fric_iter):
    """sheet sliding on the ground (mu = 0.4, 3 m/s): three steps against the oracle;
    fricIterAmt 0 re-lags the friction until converged"""
    sd = fric_scene(mu=0.4, gap=9.5e-4, v0=(3.0, 0.0, 0.0), fric_iter=fric_iter)
    _trajectory(sd, 3, v0=sd.v0)


def test_sliding_sheet_momentum_gpu(require_gpu):
    """Coulomb force in the sliding regime: sum_a m_a (v1 - v0)_a = -(mu / h) sum_k
    lambda_k B_k u_k / |u_k| after one GPU step (all contacts point-triangle, sliding)."""
    mu = 0.4
    sd = fric_scene(mu=mu, gap=9.5e-4, v0=(3.0, 0.0, 0.0))
    gpu = Simulator(sd, pcg_tol=1e-10)
    x0, _ = gpu.get_state()
    gpu.set_state(x0, sd.v0.ravel())
    C = gpu.constraint_set(x0)
    FC, FI = gpu.friction_basis(x0, C, gpu.scalars()["kappa"])
    st = gpu.step()
    assert st["status"] == 0
    x1, v1 = gpu.get_state()
    m = gpu.setup_arrays()["mass"]
    nc = sd.n_cloth
    dP = (m[:nc, None] * (v1.reshape(-1, 3)[:nc] - sd.v0[:nc])).sum(axis=0)
    h = sd.params["dt"]
    F = np.zeros(3)
    for c, f in zip(FC, FI):
        p = -c[0] - 1
        B = np.stack([f[2:5], f[5:8]], axis=1)
        u = B.T @ (x1.reshape(-1, 3)[p] - x0.reshape(-1, 3)[p])
        F += mu * f[8] * (B @ (u / np.linalg.norm(u)))
    pred = -F / h
    assert pred[0] < 0
    assert abs(dP[0] - pred[0]) <= 1e-4 * abs(pred[0])


def test_friction_codim_multiplicity(require_gpu):
    """rods and particles on the sheet (offset 1 mm): point-edge and point-point friction
    stencils with multiplicity; basis, energy and gradient against the oracle, then
    three steps of the falling codimensional mix with mu = 0.3"""
    sd = scenes.scene_codim_contact()
    sd.params.update(mu=0.3, epsv2=1e-6)
    gpu = Simulator(sd)
    orc = O.OracleScene(sd, threads=4)
    x, _ = orc.get_state()
    C = orc.constraint_set(x)
    kappa = orc.scalars()["kappa"]
    FCg, FIg = gpu.friction_basis(x, C, kappa)
    FCo, FIo = orc.friction_basis(x, C, kappa)
    assert np.array_equal(FCg, FCo)
    assert ((FCo[:, 0] < 0) & (FCo[:, 3] < -1)).any()        # multiplicity > 1 present
    for k in range(9):
        assert rel(FIg[:, k], FIo[:, k]) <= 1e-12, k
    ev = np.sqrt(1e-6) * sd.params["dt"]
    xx = x + np.random.default_rng(3).standard_normal(x.size) * ev * 3.0
    assert abs(gpu.friction_energy(xx, x) - orc.friction_energy(xx, x, FCo, FIo)) <= 1e-11 * abs(orc.friction_energy(xx, x, FCo, FIo))
    assert rel(gpu.friction_gradient(xx, x), orc.friction_gradient(xx, x, FCo, FIo)) <= 1e-11
    gpu.close()
    sd2 = scenes.scene_codim_small()
    sd2.params.update(mu=0.3, epsv2=1e-6)
    sd2.params["PNTol"] = 1e-6 * sd2.bbox_diag()
    _trajectory(sd2, 3)


def test_friction_components_gpu(require_gpu):
    """per-component coefficients (Compute_Friction_Coef): scaled normal forces against the
    oracle, then three steps of the contact stack with a different coefficient per sheet pair"""
    sd = scenes.scene_contact_stack()
    mu = np.array([[0.1, 0.2, 0.3, 0.4], [0.2, 0.5, 0.6, 0.7], [0.3, 0.6, 0.8, 0.9], [0.4, 0.7, 0.9, 1.0]])
    sd.params.update(compNodeRange=[144, 288, 432, sd.nV], muComp=mu.T.ravel(), epsv2=1e-6)
    gpu = Simulator(sd)
    orc = O.OracleScene(sd, threads=4)
    x, _ = orc.get_state()
    C = orc.constraint_set(x)
    kappa = orc.scalars()["kappa"]
    FCg, FIg = gpu.friction_basis(x, C, kappa)
    FCo, FIo = orc.friction_basis(x, C, kappa)
    assert np.array_equal(FCg, FCo)
    assert rel(FIg[:, 8], FIo[:, 8]) <= 1e-12
    gpu.close()
    sd.params["PNTol"] = 1e-6 * sd.bbox_diag()
    _trajectory(sd, 3)

"""Strain limiting on the GPU (through the C ABI) against the oracle.

Bars: validity flags equal; strain-limit energy 1e-10 relative (different SVD rounding);
elastic gradient incl. the strain-limit term 1e-10; elastic Hessian incl. the projected
strain-limit blocks 1e-9 (analytic eigen-clamp on the GPU, general eigensolver in the
oracle); steps from the GPU's own state to bars.trajectory_bar, every singular value
below the limit.
"""
import numpy as np
import pytest

import oracle as O
from bars import trajectory_bar
from cipc import Simulator
from test_oracle_strain import hammock, max_sigma

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)


@pytest.fixture(scope="module")
def stretched():
    """the hammock after 6 limited oracle steps: many triangles inside the barrier band"""
    sd = hammock(1e3)
    orc = O.OracleScene(sd)
    for _ in range(6):
        rc, _ = orc.step()
        assert rc > 0
    x, _ = orc.get_state()
    return sd, x


@pytest.mark.parametrize("limits", ["upper", "upper+lower"])
def test_strain_terms_match_oracle(require_gpu, stretched, limits):
    sd, x = stretched
    if limits == "upper+lower":
        sd = hammock(1e3)
        sd.params.update(kappa_s=(1e3, 1e3), s=(1.01, 0.95), sHat=(1.0, 1.001))   # both barriers active
    gpu = Simulator(sd)
    orc = O.OracleScene(sd)
    okg, eg = gpu.strain_limit_energy(x)
    oko, eo, gso = orc.strain_energy_grad(x)
    assert okg and oko and eo > 0
    assert abs(eg - eo) <= 1e-10 * eo
    gel, _ = orc.gradient(x, np.zeros((0, 4), np.int32), 1.0)
    assert rel(gpu.elastic_gradient(x), gel + gso) <= 1e-10
    none = np.zeros((0, 4), np.int32)
    Hg = gpu.hessian(x, none, 1.0, which=1)
    Ho = orc.hessian(x, none, 1.0, which=1)
    assert rel(Hg.toarray(), Ho.toarray()) <= 1e-9
    # beyond the limit both report an invalid potential
    X = x.reshape(-1, 3) * np.array([1.03, 1.0, 1.0])
    assert not gpu.strain_limit_energy(X.ravel())[0]
    assert not orc.strain_energy_grad(X.ravel())[0]
    gpu.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_strain_limited_trajectory(require_gpu, precision):
    """the pinned sheet sagging under gravity with the 1.01 limit: six steps, each from the
    GPU's state on both sides"""
    sd = hammock(1e3)
    gpu = Simulator(sd, pcg_tol=1e-10, precision=precision)
    orc = O.OracleScene(sd)
    probe = O.OracleScene(sd)
    for s in range(6):
        x0, v0 = gpu.get_state()
        orc.set_state(x0, v0)
        st = gpu.step()
        rc, so = orc.step()
        assert rc > 0 and st["status"] == 0, (s, st, rc)
        xg, _ = gpu.get_state()
        xo, _ = orc.get_state()
        disp = np.abs(xo - x0).max()
        assert np.abs(xg - xo).max() <= trajectory_bar(sd, disp), (s, st["newton_iters"], so)
        assert max_sigma(probe, xg, sd.nT) < 1.01
    assert max_sigma(probe, xg, sd.nT) > 1.0
    gpu.close()


def fiber_hammock():
    sd = hammock(0.0)
    sd.params.update(fiberStiffMult=(5e3, 1e3, 5e3, 5e3), inextLimit=(0.03, 0.03, 0.03))
    return sd


def test_fiber_terms_match_oracle(require_gpu):
    """anisotropic (fibre) membrane: elastic gradient and Hessian (fibre replaces the
    membrane) against the oracle at a sagged state"""
    sd = fiber_hammock()
    orc = O.OracleScene(sd)
    for _ in range(4):
        rc, _ = orc.step()
        assert rc > 0
    x, _ = orc.get_state()
    gpu = Simulator(sd)
    gel, _ = orc.gradient(x, np.zeros((0, 4), np.int32), 1.0)      # bending only (membrane replaced)
    gf = np.zeros_like(x)
    for t in range(sd.nT):
        ok, _, g9, _, _ = orc.fiber_tri(x, t)
        assert ok
        for a, v in enumerate(sd.tris[t]):
            gf[3 * v:3 * v + 3] += g9[3 * a:3 * a + 3]
    assert np.abs(gf).max() > 0
    assert rel(gpu.elastic_gradient(x), gel + gf) <= 1e-10
    none = np.zeros((0, 4), np.int32)
    assert rel(gpu.hessian(x, none, 1.0, which=1).toarray(), orc.hessian(x, none, 1.0, which=1).toarray()) <= 1e-9
    gpu.close()


def test_fiber_trajectory(require_gpu):
    """five fibre-membrane steps of the pinned sheet, each from the GPU's state, every
    Green strain component inside its limit"""
    sd = fiber_hammock()
    gpu = Simulator(sd, pcg_tol=1e-10)
    orc = O.OracleScene(sd)
    for s in range(5):
        x0, v0 = gpu.get_state()
        orc.set_state(x0, v0)
        st = gpu.step()
        rc, so = orc.step()
        assert rc > 0 and st["status"] == 0, (s, st, rc)
        xg, _ = gpu.get_state()
        xo, _ = orc.get_state()
        disp = np.abs(xo - x0).max()
        assert np.abs(xg - xo).max() <= trajectory_bar(sd, disp), (s, st["newton_iters"], so)
        e = np.array([orc.fiber_tri(xg, t)[4] for t in range(sd.nT)])
        assert np.abs(e).max() < 0.03
    assert np.abs(e).max() > 0.02          # the limit is active
    gpu.close()


def test_kappa_s_doubling_matches_oracle(require_gpu):
    """IMPLICIT_EULER.h:606-645 on both sides: with kappa_s = 1 the fourth step doubles it
    three times; each step starts from the GPU's state and kappa_s on both sides; kappa_s
    after every step equal, positions to the trajectory bar, the doubled value read back
    through cipc_get_strain_limit and kept by later steps"""
    sd = hammock(1.0)
    gpu = Simulator(sd, pcg_tol=1e-10)
    orc = O.OracleScene(sd)
    probe = O.OracleScene(sd)
    ks = []
    for s in range(5):
        x0, v0 = gpu.get_state()
        kg = gpu.strain_limit()[0]
        orc.set_state(x0, v0)
        orc.set_strain_limit(kg, sd.params["s"], sd.params["sHat"])
        st = gpu.step()
        rc, so = orc.step()
        assert rc > 0 and st["status"] == 0, (s, st, rc)
        assert st["kappa_s"] == so["kappa_s"] == gpu.strain_limit()[0][0], (s, st["kappa_s"], so["kappa_s"])
        assert st["kappa_s_doublings"] == so["kappa_s_doublings"]
        ks.append(st["kappa_s"])
        xg, _ = gpu.get_state()
        xo, _ = orc.get_state()
        disp = np.abs(xo - x0).max()
        assert np.abs(xg - xo).max() <= trajectory_bar(sd, disp), (s, st["newton_iters"], so)
        assert max_sigma(probe, xg, sd.nT) < 1.01
    assert ks == [1.0, 1.0, 1.0, 8.0, 8.0], ks
    gpu.close()

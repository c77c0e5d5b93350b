"""Full-size and edge-case parity of the HIP path (through the C ABI) against the oracle.

* The 320k-triangle stack (four 201x201 sheets over the sphere) stepped on the GPU until
  the sheets lie on the sphere: every step intersection free, and at the contact state
  the constraint set and distances bit-exact against the oracle, barrier energy and
  gradient to summation order, the CCD step no later than the oracle's.
* Empty inputs: a state without contacts gives an empty constraint set, zero barrier
  energy and gradient, and no constraint bound on the CCD step.
* Closed-form edge-edge CCD cases on the GPU (crossing and parallel edges).
"""
import numpy as np
import pytest

import oracle as O
from cipc import Simulator, scenes

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)


def test_c2_contact_state(require_gpu):
    sd = scenes.scene_c2()
    gpu = Simulator(sd, precision="fp32")
    reached = False
    for s in range(40):
        st = gpu.step()
        assert st["status"] == 0 and st["newton_iters"] > 0, (s, st)
        print("step", s, st["contacts"], st["newton_iters"], st["pcg_iters"], flush=True)
        if st["contacts"] > 0:
            x = gpu.get_state()[0]
            assert gpu.min_dist2(x, gpu.constraint_set(x))[1] > 0
        if st["contacts"] >= 5000:
            reached = True
            break
    assert reached, "fewer than 5000 contacts after 40 steps"
    x, _ = gpu.get_state()
    orc = O.OracleScene(sd, threads=8)
    Cg = gpu.constraint_set(x)
    Co = orc.constraint_set(x)
    print("constraints", len(Cg), flush=True)
    assert len(Co) > 1000
    assert np.array_equal(Cg, Co)
    d2g, mg = gpu.min_dist2(x, Co)
    d2o = orc.dist2_list(x, Co)
    assert np.array_equal(d2g, d2o)
    assert mg == d2o.min() - sd.params["xi"] ** 2 and mg > 0
    kappa = gpu.scalars()["kappa"]
    eo, bad = orc.energy(x, None, Co, kappa)
    assert bad == 0
    assert abs(gpu.barrier_energy(x, Co, kappa) - eo[3]) <= 1e-11 * abs(eo[3])
    _, gbo = orc.gradient(x, Co, kappa)
    assert rel(gpu.barrier_gradient(x, Co, kappa), gbo) <= 1e-11
    rng = np.random.default_rng(11)
    dbc = np.zeros(sd.nV, bool)
    dbc[sd.dbc_idx] = True
    for scale in (1e-4, 1e-3):
        p = rng.standard_normal(x.size) * scale
        p.reshape(-1, 3)[dbc] = 0
        ao, clo = orc.ccd(x, p, 1.0)
        ag, clg = gpu.ccd_stepsize(x, p, 1.0)
        print("ccd", scale, ag, ao, clg, flush=True)
        assert clg == clo
        assert ao * (1 - 1e-9) <= ag <= ao      # never later than the FP64 oracle
    gpu.close()


def test_empty_constraint_inputs(require_gpu):
    sd = scenes.scene_c1()
    gpu = Simulator(sd)
    x, _ = gpu.get_state()
    C = gpu.constraint_set(x)
    assert C.shape == (0, 4)
    kappa = gpu.scalars()["kappa"]
    assert gpu.barrier_energy(x, C, kappa) == 0.0
    assert not gpu.barrier_gradient(x, C, kappa).any()
    d2, m = gpu.min_dist2(x, C)
    assert d2.size == 0 and m > 1e200
    assert gpu.hessian(x, C, kappa, which=2).nnz == 0
    # a small free-fall direction: no constraint limits the step, only the hash clamp
    p = np.zeros_like(x).reshape(-1, 3)
    p[: sd.n_cloth, 1] = -1e-3
    orc = O.OracleScene(sd)
    ag, clg = gpu.ccd_stepsize(x, p.ravel(), 1.0)
    ao, clo = orc.ccd(x, p.ravel(), 1.0)
    assert (ag, clg) == (ao, clo) == (1.0, 1.0)
    gpu.close()


@pytest.mark.parametrize("layout", ["crossing", "parallel"])
def test_ccd_edge_edge_closed_form_gpu(require_gpu, layout):
    """Same cases as the oracle's closed-form edge-edge test: exact time of impact d0 / v,
    the GPU step within [0.85, 1] of it, never later than the oracle's and equal to it to
    rounding."""
    d0 = 0.01
    A = [[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, -1.0, 0.0]]
    if layout == "crossing":
        B = [[0.0, d0, -1.0], [0.0, d0, 1.0], [0.0, d0 + 1.0, 0.0]]
    else:
        B = [[-0.5, d0, 0.0], [0.5, d0, 0.0], [0.0, d0 + 1.0, 1.0]]
    X = np.array(A + B)
    T = np.array([[0, 1, 2], [3, 4, 5]], dtype=np.int32)
    fixed = np.array([0, 1, 2], dtype=np.int32)
    sd = scenes.SceneData(X, T, fixed, X[fixed].copy(), dict(scenes.DEFAULT_PARAMS, PNTol=1e-3, xi=0.0))
    gpu = Simulator(sd)
    orc = O.OracleScene(sd)
    x = X.ravel().copy()
    for v in (0.02, 0.2):
        p = np.zeros_like(x).reshape(-1, 3)
        p[3:, 1] = -v
        ag, clg = gpu.ccd_stepsize(x, p.ravel(), 1.0)
        ao, clo = orc.ccd(x, p.ravel(), 1.0)
        toi = d0 / v
        assert clg == clo
        assert 0.85 * toi <= ag <= toi
        assert ao * (1 - 1e-12) <= ag <= ao      # never later than the FP64 oracle
    gpu.close()

"""The benchmark configurations on the GPU (through the C ABI) against the oracle.

* C1, the 41x41 drape (1,681 cloth vertices over the icosphere and the ground): all 100
  steps, each started from the GPU's own state (positions, velocities, barrier
  stiffness) on both sides and compared with the oracle's step to bars.trajectory_bar;
  every state intersection free; at least 20 of the steps in contact.
* C4 on the same steps: the mixed-precision learned form (FP32 blocks and vectors, FP64
  refinement, learned factor with block-Jacobi fallback, PCG to 1e-5) meets the same bar
  from the same states; its PCG iterations and fallbacks are set against block-Jacobi
  PCG at the same tolerance.
* C5: the 256-scene batch (triangle counts log-uniform in [1e3, 1e5]) in the learned form:
  every scene's step succeeds and ends intersection free; a sample of scenes stepped one
  by one from the same state gives the same bits.
* C3 at full size (C1 + 20,000 rods + 100,000 particles, offset 1 mm): stepped until the
  rods and particles land on the cloth; at that state the constraint set (with its
  point-edge / point-point codimensional entries) and distances are bit-exact against
  the oracle, barrier energy and gradient agree to summation order, the CCD step is never
  later than the oracle's.
"""
import numpy as np
import pytest

import oracle as O
from bars import trajectory_bar
from cipc import BatchStepper, Simulator, precond, scenes

pytestmark = pytest.mark.gpu

C1_STEPS = 100
C1_BLOCK = 25


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def c1_chain():
    sd = scenes.scene_c1()
    ref = Simulator(sd, pcg_tol=1e-10)                     # the chain: FP64 block Jacobi, tight solves
    learned = Simulator(sd, pcg_tol=1e-5, precision="fp32")
    learned.precond_set_weights(precond.load())
    learned.set_preconditioner("learned", budget=2500)
    jacobi = Simulator(sd, pcg_tol=1e-5)                   # block Jacobi at the learned form's tolerance
    orc = O.OracleScene(sd, threads=16)
    chain = dict(sd=sd, ref=ref, learned=learned, jacobi=jacobi, orc=orc, step=0, contact_steps=0,
                 pcg_learned=0, pcg_jacobi=0, newton_learned=0)
    yield chain
    for k in ("ref", "learned", "jacobi"):
        chain[k].close()


@pytest.mark.parametrize("block", range(C1_STEPS // C1_BLOCK))
def test_c1_drape_and_c4_learned_form_match_oracle(require_gpu, c1_chain, block):
    ch = c1_chain
    sd, ref, learned, jacobi, orc = ch["sd"], ch["ref"], ch["learned"], ch["jacobi"], ch["orc"]
    assert ch["step"] == block * C1_BLOCK, "an earlier block failed"
    for k in range(C1_BLOCK):
        s = ch["step"]
        x0, v0 = ref.get_state()
        kappa = ref.scalars()["kappa"]
        for sim in (orc, learned, jacobi):
            sim.set_state(x0, v0)
            sim.set_kappa(kappa)
        st = ref.step()
        rc, so = orc.step()
        sl = learned.step()
        sj = jacobi.step()
        assert rc > 0 and st["status"] == 0 and sl["status"] == 0 and sj["status"] == 0, (s, st, rc, sl, sj)
        xo, _ = orc.get_state()
        bar = trajectory_bar(sd, np.abs(xo - x0).max())
        xg, _ = ref.get_state()
        xl, _ = learned.get_state()
        assert np.abs(xg - xo).max() <= bar, (s, st["newton_iters"], so["newton_iters"])
        assert np.abs(xl - xo).max() <= bar, (s, sl["newton_iters"], so["newton_iters"])
        C = ref.constraint_set(xg)
        if len(C):
            assert ref.min_dist2(xg, C)[1] > 0
            ch["contact_steps"] += 1
            ch["pcg_learned"] += sl["pcg_iters"]
            ch["pcg_jacobi"] += sj["pcg_iters"]
            ch["newton_learned"] += sl["newton_iters"]
        print("C1 step %d: %d contacts, Newton %d (oracle %d), PCG learned %d / Jacobi %d" %
              (s + 1, len(C), st["newton_iters"], so["newton_iters"], sl["pcg_iters"], sj["pcg_iters"]), flush=True)
        ch["step"] += 1
    if ch["step"] == C1_STEPS:
        info = learned.precond_info()
        print("C4 on C1, %d contact steps: PCG iterations learned %d vs block Jacobi %d; learned solves %d, fallbacks %d" %
              (ch["contact_steps"], ch["pcg_learned"], ch["pcg_jacobi"], info["solves"], info["fallbacks"]), flush=True)
        assert ch["contact_steps"] >= 20
        assert ch["pcg_learned"] < ch["pcg_jacobi"]
        assert info["fallbacks"] <= info["solves"] // 10


def test_c5_batch_learned_form(require_gpu):
    sds = scenes.scene_c5(count=256)
    w = precond.load()
    sims = []
    for sd in sds:
        sim = Simulator(sd, pcg_tol=1e-5, precision="fp32")
        sim.precond_set_weights(w)
        sim.set_preconditioner("learned", budget=2500)
        sims.append(sim)
    sample = [0, 37, 101, 255]
    solo = []
    for i in sample:
        s = Simulator(sds[i], pcg_tol=1e-5, precision="fp32")
        s.precond_set_weights(w)
        s.set_preconditioner("learned", budget=2500)
        solo.append(s)
    batch = BatchStepper(sims, threads=16)
    for k in range(3):
        st, ms = batch.step()
        assert all(s["status"] == 0 for s in st)
        for s in solo:
            s.step()
        print("C5 batch step %d: p50 %.1f ms, p99 %.1f ms" % (k + 1, np.percentile(ms, 50), np.percentile(ms, 99)), flush=True)
    for sim in sims:
        x, _ = sim.get_state()
        C = sim.constraint_set(x)
        if len(C):
            assert sim.min_dist2(x, C)[1] > 0
    for i, s in zip(sample, solo):
        xa, va = sims[i].get_state()
        xb, vb = s.get_state()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), i
    for s in sims + solo:
        s.close()


def test_c3_full_size_codim_contact(require_gpu):
    sd = scenes.scene_c3()
    gpu = Simulator(sd)
    codim0 = gpu.codimBN0
    for s in range(40):
        st = gpu.step()
        assert st["status"] == 0, (s, st)
        x = gpu.get_state()[0]
        C = gpu.constraint_set(x)
        if len(C):
            assert gpu.min_dist2(x, C)[1] > 0
        # entries whose point is a rod node or a particle (boundary-node slot >= codim0)
        pt = -C[C[:, 0] < 0, 0] - 1
        codim = int(np.isin(pt, gpu.codim_setup()["bnode"][codim0:]).sum())
        print("C3 step %d: %d contacts (%d with a codimensional point), Newton %d" % (s + 1, len(C), codim, st["newton_iters"]),
              flush=True)
        if codim >= 1000:
            break
    assert codim >= 1000, "the rods and particles did not reach the cloth"
    orc = O.OracleScene(sd, threads=16)
    Co = orc.constraint_set(x)
    assert np.array_equal(C, Co)
    assert ((Co[:, 0] < 0) & (Co[:, 3] < 0)).any()            # point-edge / point-point entries
    d2g, mg = gpu.min_dist2(x, Co)
    d2o = orc.dist2_list(x, Co)
    assert np.array_equal(d2g, d2o)
    kappa = gpu.scalars()["kappa"]
    eo, bad = orc.energy(x, None, Co, kappa)
    assert bad == 0
    assert abs(gpu.barrier_energy(x, Co, kappa) - eo[3]) <= 1e-11 * abs(eo[3])
    _, gbo = orc.gradient(x, Co, kappa)
    assert rel(gpu.barrier_gradient(x, Co, kappa), gbo) <= 1e-11
    rng = np.random.default_rng(3)
    dbc = np.zeros(sd.nV, bool)
    dbc[sd.dbc_idx] = True
    for scale in (1e-4, 1e-3):
        p = rng.standard_normal(x.size) * scale
        p.reshape(-1, 3)[dbc] = 0
        ao, clo = orc.ccd(x, p, 1.0)
        ag, clg = gpu.ccd_stepsize(x, p, 1.0)
        assert clg == clo
        assert ao * (1 - 1e-9) <= ag <= ao
    gpu.close()

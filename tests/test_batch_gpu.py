"""Batched stepping (cipc_batch_step): scenes stepped together by several host threads
equal the same scenes stepped one by one, bit for bit (every kernel is deterministic and
a scene's work stays on its own stream)."""
import numpy as np
import pytest

from cipc import BatchStepper, Simulator, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("features", ["plain", "friction+strain-limit"])
def test_batch_equals_sequential(require_gpu, features):
    sds = scenes.scene_c5(seed=7, count=6, tri_min=500, tri_max=3000)
    for sd in sds:      # start just above the sphere so that contact happens within the steps
        top = sd.X[sd.n_cloth:, 1].max()
        sd.X[:sd.n_cloth, 1] += top + 0.004 - sd.X[:sd.n_cloth, 1].min()
        if features != "plain":
            sd.params.update(mu=0.3, epsv2=1e-6, kappa_s=(1e3, 0.0), s=(1.05, 0.0), sHat=(1.0, 1.0))
    a = [Simulator(sd) for sd in sds]
    b = [Simulator(sd) for sd in sds]
    batch = BatchStepper(a, threads=4)
    for _ in range(4):
        st, ms = batch.step()
        for s in b:
            s.step()
        assert len(ms) == len(sds) and np.all(ms > 0)
    for sa, sb in zip(a, b):
        xa, va = sa.get_state()
        xb, vb = sb.get_state()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    assert sum(s["contacts"] for s in st) > 0

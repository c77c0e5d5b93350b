"""The allocation-free barrier energy (cipc_barrier_energy_noalloc) against the allocating
entry point (cipc_barrier_energy) through the C ABI.

The no-alloc call stages positions, constraint rows, block partials, the sum and the
validity flags in the library's static per-GPU scratch array; it runs the same kernel
arithmetic in the same block partition, so the energies must agree bit for bit.  Two
calls in a row (the second with a smaller input, so stale contents of the shared array
would show), then a second context (another scene, fewer vertices) on the same GPU, the
error paths (a zero-distance constraint, an out-of-range entry, an input beyond the
array's capacity) and an empty set.
"""
import numpy as np
import pytest

from cipc import CipcError, Simulator, scenes

pytestmark = pytest.mark.gpu


def test_noalloc_energy_equals_allocating_entry(require_gpu):
    sd = scenes.scene_contact_stack(n=10)
    sim = Simulator(sd)
    x0, _ = sim.get_state()
    C = sim.constraint_set(x0)
    assert len(C) > 1000
    kappa = sim.scalars()["kappa"]
    e_ref = sim.barrier_energy(x0, C, kappa)
    e_na = sim.barrier_energy_noalloc(x0, C, kappa)
    print("noalloc call 1: %d constraints, E %.17g vs %.17g" % (len(C), e_na, e_ref), flush=True)
    assert e_na == e_ref and e_ref > 0

    # second call: fewer constraints and moved positions (the first call's larger contents
    # are still in the shared array)
    rng = np.random.default_rng(11)
    x1 = x0 + rng.uniform(-1e-5, 1e-5, x0.shape) * 0.01
    C1 = C[::3].copy()
    e_ref1 = sim.barrier_energy(x1, C1, kappa)
    e_na1 = sim.barrier_energy_noalloc(x1, C1, kappa)
    print("noalloc call 2: %d constraints, E %.17g vs %.17g" % (len(C1), e_na1, e_ref1), flush=True)
    assert e_na1 == e_ref1 and e_ref1 != e_ref

    # a second context on the same GPU shares the array
    sd2 = scenes.scene_cloth_on_ground(n=12)
    sim2 = Simulator(sd2)
    x2, _ = sim2.get_state()
    C2 = sim2.constraint_set(x2)
    assert len(C2) > 0
    k2 = sim2.scalars()["kappa"]
    assert sim2.nV < sim.nV
    assert sim2.barrier_energy_noalloc(x2, C2, k2) == sim2.barrier_energy(x2, C2, k2)
    # and the first context still gets its own result afterwards
    assert sim.barrier_energy_noalloc(x0, C, kappa) == e_ref

    # empty set
    empty = np.zeros((0, 4), np.int32)
    assert sim.barrier_energy_noalloc(x0, empty, kappa) == 0.0

    # a point-point constraint of a vertex with itself: distance 0, as the allocating call
    bad = np.array([[-1, 0, -1, -1]], np.int32)
    with pytest.raises(CipcError) as e1:
        sim.barrier_energy(x0, bad, kappa)
    with pytest.raises(CipcError) as e2:
        sim.barrier_energy_noalloc(x0, bad, kappa)
    assert e1.value.code == e2.value.code == -3

    # an entry outside [-n_vertices, n_vertices) is refused on the host
    oob = np.array([[-1, sim.nV, -1, -1]], np.int32)
    with pytest.raises(CipcError) as e3:
        sim.barrier_energy_noalloc(x0, oob, kappa)
    assert e3.value.code == -1

    # more than the array holds: refused before any device work, the context stays usable
    big = np.tile(C[:1], (3_000_000, 1))
    with pytest.raises(CipcError) as e4:
        sim.barrier_energy_noalloc(x0, big, kappa)
    assert e4.value.code == -7
    assert sim.barrier_energy_noalloc(x1, C1, kappa) == e_ref1
    sim2.close()
    sim.close()

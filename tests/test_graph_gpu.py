"""Graph-captured PCG chunks (cipc_set_graphs) against the same chunks launched kernel by
kernel: identical bits for the linear solve over repeated replays (with the caller's own
eager work and synchronisations between them), for the learned and the block-Jacobi
preconditioners and the three precision forms, and for whole contact steps."""
import numpy as np
import pytest

from cipc import Simulator, precond, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,learned", [("fp64", False), ("mixed", False), ("fp32", False), ("fp32", True)])
def test_graph_solves_equal_launched_solves(require_gpu, precision, learned):
    sd = scenes.scene_contact_stack()
    sims = []
    for graphs in (False, True):
        s = Simulator(sd, precision=precision)
        s.set_graphs(graphs)
        if learned:
            s.precond_set_weights(precond.load())
            s.set_preconditioner("learned", budget=100000)
        sims.append(s)
    x, _ = sims[0].get_state()
    C = sims[0].constraint_set(x)
    kappa = sims[0].scalars()["kappa"]
    rng = np.random.default_rng(2)
    for rep in range(4):
        b = rng.standard_normal(x.size)
        b.reshape(-1, 3)[sd.dbc_idx] = 0
        out = [s.linear_solve(x, C, kappa, b, tol=1e-9) for s in sims]
        assert out[0][1] == out[1][1] and out[0][1] > 40, (rep, out[0][1], out[1][1])
        assert np.array_equal(out[0][0], out[1][0]), rep
        assert out[0][2] == out[1][2]
    info = sims[1].solver_info()
    assert info["graph_captures"] >= 1 and info["graph_launches"] >= 4
    assert sims[0].solver_info()["graph_launches"] == 0
    for s in sims:
        s.close()


def test_graph_steps_equal_launched_steps(require_gpu):
    sd = scenes.scene_contact_stack()
    a, b = Simulator(sd, pcg_tol=1e-8), Simulator(sd, pcg_tol=1e-8)
    a.set_graphs(False)
    for _ in range(3):
        sa, sb = a.step(), b.step()
        assert sa["newton_iters"] == sb["newton_iters"] and sa["pcg_iters"] == sb["pcg_iters"]
    xa, va = a.get_state()
    xb, vb = b.get_state()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    assert b.solver_info()["graph_launches"] > 0
    a.close()
    b.close()

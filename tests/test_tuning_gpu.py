"""Autotuner (cipc_autotune): it settles within the first steps (re-tuning at most twice
as contacts change the matrix) on shapes from its candidate sets, leaves the constraint
set unchanged (the grid only affects speed), and each step of the tuned simulation
agrees with an FP64 oracle step from the same state to the trajectory bar of the parity
tests (the launch shapes change only the order of the PCG reductions)."""
import numpy as np
import pytest

import oracle as O
from bars import trajectory_bar
from cipc import Simulator, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotune_settles_and_keeps_results(require_gpu, precision):
    sd = scenes.scene_contact_stack()
    sd.params["PNTol"] = 1e-6 * sd.bbox_diag()
    a = Simulator(sd, pcg_tol=1e-10, precision=precision)
    b = Simulator(sd, pcg_tol=1e-10, precision=precision)
    orc = O.OracleScene(sd, threads=4)
    a.autotune()
    assert a.tuning()["tuning_pcg"] and a.tuning()["tuning_cell"]
    x0, _ = a.get_state()
    C0 = b.constraint_set(x0)
    for s in range(12):
        xs, vs = a.get_state()
        orc.set_state(xs, vs)
        orc.set_kappa(a.scalars()["kappa"])
        assert a.step()["status"] == 0
        rc, _ = orc.step()
        assert rc > 0
        xa, _ = a.get_state()
        xo, _ = orc.get_state()
        disp = np.abs(xo - xs).max()
        assert np.abs(xa - xo).max() <= trajectory_bar(sd, disp), s
    t = a.tuning()
    assert not t["tuning_pcg"] and not t["tuning_cell"]
    assert t["spmv_grid"] in (1024, 2048) and t["spmv_lanes"] in (4, 8, 16) and t["spmv_unroll"] in (1, 2)
    assert t["cell_mult"] in (0.5, 0.7, 1.0, 1.4, 2.0)
    assert np.array_equal(a.constraint_set(x0), C0)

"""Autotuner (cipc_autotune): it settles within the first steps on shapes from its
candidate sets, leaves the constraint set unchanged (the grid only affects speed), and
the tuned simulation follows the untuned one to the trajectory bar of the parity tests."""
import numpy as np
import pytest

from cipc import Simulator, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotune_settles_and_keeps_results(require_gpu, precision):
    sd = scenes.scene_contact_stack()
    sd.params["PNTol"] = 1e-6 * sd.bbox_diag()
    a = Simulator(sd, pcg_tol=1e-10, precision=precision)
    b = Simulator(sd, pcg_tol=1e-10, precision=precision)
    a.autotune()
    assert a.tuning()["tuning_pcg"] and a.tuning()["tuning_cell"]
    x0, _ = a.get_state()
    C0 = b.constraint_set(x0)
    for _ in range(6):
        a.step()
        b.step()
    t = a.tuning()
    assert not t["tuning_pcg"] and not t["tuning_cell"]
    assert t["spmv_grid"] in (1024, 2048) and t["cell_mult"] in (0.5, 0.7, 1.0, 1.4, 2.0)
    assert np.array_equal(a.constraint_set(x0), C0)
    xa, _ = a.get_state()
    xb, _ = b.get_state()
    disp = np.abs(xb - sd.X.ravel()).max()
    assert np.abs(xa - xb).max() <= 1e-6 * max(disp, 1e-3)

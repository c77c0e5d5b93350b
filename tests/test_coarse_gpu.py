"""Two-level correction of the learned preconditioner (coarse.hip) through the C ABI.

Bars: every solve meets the FP64 residual bar of the oracle's own assembled matrix (the
coarse level changes the preconditioner, never the stopping rule); with the trained
weights it needs fewer PCG iterations than the learned factor alone; the Galerkin matrix
is summed in a fixed order, so repeated solves and graph-replayed chunks give the same
bits; steps with it stay on the oracle's trajectory (bars.trajectory_bar).
"""
import numpy as np
import pytest

import oracle as O
from bars import trajectory_bar
from cipc import Simulator, precond, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ground(require_gpu):
    sd = scenes.scene_cloth_on_ground()
    orc = O.OracleScene(sd, threads=4)
    for _ in range(2):
        rc, _ = orc.step()
        assert rc > 0
    x = orc.get_state()[0].copy()
    kappa = orc.scalars()["kappa"]
    C = orc.constraint_set(x)
    H = orc.hessian(x, C, kappa, which=3)
    b = np.random.default_rng(12).standard_normal(x.size)
    b.reshape(-1, 3)[sd.dbc_idx] = 0
    return sd, x, C, kappa, H, b


def learned_sim(sd, precision, coarse, graphs=True, bf16=False):
    sim = Simulator(sd, precision=precision)
    sim.precond_set_weights(precond.load())
    sim.set_preconditioner("learned", budget=20000)
    sim.set_coarse(coarse, bf16=bf16)
    sim.set_graphs(graphs)
    return sim


@pytest.mark.parametrize("precision,storage", [("fp64", "fp32"), ("mixed", "fp32"), ("fp32", "fp32"), ("mixed", "bf16"),
                                               ("fp32", "bf16")])
def test_coarse_solve_meets_bar_and_cuts_iterations(ground, precision, storage):
    sd, x, C, kappa, H, b = ground
    its = {}
    for coarse in (False, True):
        sim = learned_sim(sd, precision, coarse, bf16=storage == "bf16")
        sol, it, rr = sim.linear_solve(x, C, kappa, b, tol=1e-8)
        assert sim.precond_info()["fallback"] == 0
        assert rr <= 1e-8
        assert np.linalg.norm(H @ sol - b) <= 2e-8 * np.linalg.norm(b)
        if coarse:
            info = sim.coarse_info()
            print(precision, info)
            assert info["ready"] and info["aggregates"] > 0 and info["unknowns"] == 6 * info["aggregates"]
        its[coarse] = it
        sim.close()
    print(precision, "learned %d, learned + coarse %d PCG iterations" % (its[False], its[True]))
    assert its[True] < 0.7 * its[False]


@pytest.mark.parametrize("storage", ["fp32", "bf16"])
def test_coarse_solves_repeat_bits_with_and_without_graphs(ground, storage):
    sd, x, C, kappa, H, b = ground
    sols = []
    for graphs in (True, True, False):
        sim = learned_sim(sd, "mixed", True, graphs, bf16=storage == "bf16")
        sol, it, rr = sim.linear_solve(x, C, kappa, b, tol=1e-10)
        sols.append((sol, it))
        sim.close()
    assert sols[0][1] == sols[1][1] == sols[2][1]
    assert np.array_equal(sols[0][0], sols[1][0])
    assert np.array_equal(sols[0][0], sols[2][0])


def test_coarse_step_trajectory(require_gpu):
    """Four steps of the contact stack with the learned + coarse mixed form, each from the
    GPU's state on the oracle too (as tests/test_gpu_parity.py::test_step_trajectory)."""
    sd = scenes.scene_contact_stack()
    sd.params["PNTol"] = 1e-6 * sd.bbox_diag()
    gpu = Simulator(sd, pcg_tol=1e-10, precision="mixed")
    gpu.precond_set_weights(precond.load())
    gpu.set_preconditioner("learned", budget=20000)
    gpu.set_coarse(True)
    orc = O.OracleScene(sd, threads=4)
    for s in range(4):
        x0, v0 = gpu.get_state()
        orc.set_state(x0, v0)
        orc.set_kappa(gpu.scalars()["kappa"])
        st = gpu.step()
        rc, so = orc.step()
        assert rc > 0 and st["status"] == 0
        xg, _ = gpu.get_state()
        xo, _ = orc.get_state()
        disp = np.abs(xo - x0).max()
        assert np.abs(xg - xo).max() <= trajectory_bar(sd, disp), (s, st["newton_iters"], so)
    assert gpu.coarse_info()["setups"] > 0


@pytest.mark.parametrize("max_unknowns", [0, 12])
def test_coarse_with_rods_particles_and_a_tight_bound(require_gpu, max_unknowns):
    """Rods, particles and a sheet: particles (no elastic neighbours) stay outside the coarse
    space, collinear rod aggregates lose their degenerate rotation; with a bound of 12
    unknowns only the two largest aggregates stay.  The solve meets the oracle bar."""
    sd = scenes.scene_codim_contact()
    orc = O.OracleScene(sd, threads=4)
    x = orc.get_state()[0].copy()
    kappa = orc.scalars()["kappa"]
    C = orc.constraint_set(x)
    H = orc.hessian(x, C, kappa, which=3)
    b = np.random.default_rng(21).standard_normal(x.size)
    b.reshape(-1, 3)[sd.dbc_idx] = 0
    sim = Simulator(sd, precision="fp64")
    sim.precond_set_weights(precond.load())
    sim.set_preconditioner("learned", budget=20000)
    sim.set_coarse(True, max_unknowns=max_unknowns)
    sol, it, rr = sim.linear_solve(x, C, kappa, b, tol=1e-8)
    info = sim.coarse_info()
    print(max_unknowns, info, it)
    assert info["ready"] and info["aggregates"] > 0
    if max_unknowns:
        assert info["unknowns"] <= max_unknowns
    assert rr <= 1e-8
    assert np.linalg.norm(H @ sol - b) <= 2e-8 * np.linalg.norm(b)

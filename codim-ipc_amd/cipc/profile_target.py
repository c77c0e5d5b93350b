"""A short, steady workload for profilers: each hot stage of the Newton step once at a
saved contact state, with a progress line after every stage.

    python -m cipc.profile_target STATE.npz [--workload c2] [--precision mixed]
                                           [--no-learned] [--coarse MAX_UNKNOWNS] [--repeat N]

STATE.npz is written by `bench.py --save-state` (x, v, kappa, step).  Stages, in order:
constraint set (Compute_Constraint_Set), barrier energy and gradient, the projected
Hessian (elastic + barrier: membrane, hinge and barrier PSD projections, assembly), the
learned factor (the network's forward pass on the matrix cores), one Newton linear solve
(right-hand side: the step's first one, -grad of the incremental potential at x) in the
given precision form with the learned or block-Jacobi preconditioner, and the CCD step
size along that solve's direction.  Launch shapes are the library defaults (no tuning),
graph capture is off (every kernel is its own dispatch, as tracers expect).  The solve
also reports the PCG SpMV's mean dispatch time (event-timed on every 8th launch) and its
algorithmic bytes per launch at this state.
"""
import argparse
import sys
import time

import numpy as np

from . import precond, scenes
from .sim import Simulator


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("state")
    ap.add_argument("--workload", default="c2", choices=["c1", "c2"])
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed", "fp32"])
    ap.add_argument("--no-learned", dest="learned", action="store_false")
    ap.add_argument("--repeat", type=int, default=1)
    ap.add_argument("--coarse", type=int, default=-1, metavar="MAX_UNKNOWNS",
                    help="two-level correction of the learned preconditioner with at most this many coarse unknowns "
                         "(0: the library's bound)")
    a = ap.parse_args()
    t0 = time.time()

    def log(m):
        print("[%6.1fs] %s" % (time.time() - t0, m), flush=True)

    sd = {"c1": scenes.scene_c1, "c2": scenes.scene_c2}[a.workload]()
    with np.load(a.state, allow_pickle=False) as z:
        x, v, kappa = z["x"].copy(), z["v"].copy(), float(z["kappa"])
    sim = Simulator(sd, pcg_tol=1e-5, precision=a.precision)
    sim.set_graphs(False)
    if a.learned:
        sim.precond_set_weights(precond.load())
        sim.set_preconditioner("learned", budget=20000)
        if a.coarse >= 0:
            sim.set_coarse(True, max_unknowns=a.coarse)
    sim.set_state(x, v)
    sim.set_kappa(kappa)
    mass = sim.setup_arrays()["mass"]
    dbc = np.zeros(sd.nV, bool)
    dbc[sd.dbc_idx] = True
    h = sd.params["dt"]
    grav = np.asarray(sd.params["gravity"], float)
    for r in range(a.repeat):
        C = sim.constraint_set(x)
        log("constraint set: %d constraints" % len(C))
        e = sim.barrier_energy(x, C, kappa)
        gb = sim.barrier_gradient(x, C, kappa)
        log("barrier energy %.6e and gradient" % e)
        H = sim.hessian(x, C, kappa, which=3)
        log("projected Hessian: %d stored entries" % H.nnz)
        if a.learned:
            G = sim.precond_factor(x, C, kappa, mfma=True)
            log("learned factor: %d blocks" % G.shape[1])
        xt = (x.reshape(-1, 3) + h * v.reshape(-1, 3) + h * h * np.where(mass[:, None] > 0, grav[None, :], 0.0)).ravel()
        b = -(sim.elastic_gradient(x) + gb + np.repeat(mass, 3) * (x - xt))
        b.reshape(-1, 3)[dbc] = 0
        sim.set_profiling(True)
        ts = time.time()
        p, it, rel = sim.linear_solve(x, C, kappa, b, tol=1e-5)
        ts = time.time() - ts
        pr = sim.profile()
        sim.set_profiling(False)
        form = ("learned + coarse" if a.coarse >= 0 else "learned") if a.learned else "block Jacobi"
        log("linear solve (%s, %s): %d PCG iterations, relative residual %.2e, %.1f ms with assembly" % (a.precision, form, it, rel,
                                                                                                       1e3 * ts))
        if a.coarse >= 0:
            log("coarse level: %s" % sim.coarse_info())
        if pr["launches"]:
            log("PCG SpMV: %d timed launches, %.2f us mean, %d algorithmic bytes per launch (%d block rows, %d stored blocks)"
                % (pr["launches"], 1e3 * pr["total_ms"] / pr["launches"], pr["bytes_per_launch"], pr["block_rows"], pr["nnzb"]))
        alpha, _ = sim.ccd_stepsize(x, p, 1.0)
        log("CCD step size %.6g" % alpha)
    sim.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Training of the learned preconditioner on the Newton systems the simulator produces.

Training data: cloth scenes of the two benchmark families -- the 41x41 single sheet over
the icosphere and ground (2.5 cm spacing) and two-sheet 61x61 crops at the 320k stack's
5 mm spacing -- with seeded drop heights and Young's moduli drawn log-uniform in
[1e5, 1e7] Pa.  Each scene is stepped by the GPU simulator; after every step the state,
its active constraint set and kappa define one Newton system (the Hessian the step's last
Newton iteration solved).  The network is trained on all systems in rotation with Adam
on the Kaporin loss (cipc_precond_train_step); weights and optimiser state move between
scenes through get/set state.  The result is saved as weights/precond.npz.

Usage:  python -m cipc.precond [--steps N] [--out PATH]
"""
import argparse
import math
import os
import time

import numpy as np

from . import scenes
from .sim import Simulator

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "precond.npz")


def _crop(sd, rng):
    """a 0.3 m crop of a 61x61 sheet stack: the 320k stack's 5 mm spacing"""
    X = sd.X.copy()
    X[:sd.n_cloth, 0] *= 0.3
    X[:sd.n_cloth, 2] *= 0.3
    sd2 = scenes.SceneData(X, sd.tris, sd.dbc_idx, sd.dbc_target, sd.params, sd.name)
    sd2.params["E"] = float(10 ** rng.uniform(5, 7))
    sd2.params["PNTol"] = 1e-2 * sd2.bbox_diag()
    sd2.n_cloth = sd.n_cloth
    return sd2


def training_scenes(seed=0, n_coarse=6, n_fine=3, n_stack=0, stack_steps=24):
    """n_coarse 41x41 sheets, n_fine two-sheet 61x61 crops, n_stack four-sheet 61x61 crops
    stepped `stack_steps` steps (deep self-contact, as in the 320k stack once it drapes)"""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_coarse):
        sd = scenes.cloth_on_sphere(sheets=1, n=41, y0=float(rng.uniform(0.205, 0.26)), seed=1000 + k, name="train_c41_%d" % k)
        sd.params["E"] = float(10 ** rng.uniform(5, 7))
        out.append(sd)
    for k in range(n_fine):
        sd = scenes.cloth_on_sphere(sheets=2, n=61, y0=float(rng.uniform(0.203, 0.23)), dy=0.005, seed=2000 + k, name="train_c61_%d" % k)
        out.append(_crop(sd, rng))
    for k in range(n_stack):
        sd = scenes.cloth_on_sphere(sheets=4, n=61, y0=float(rng.uniform(0.201, 0.21)), dy=0.005, seed=3000 + k, name="train_s61_%d" % k)
        sd = _crop(sd, rng)
        sd.train_steps = stack_steps
        out.append(sd)
    return out


def collect(sd, steps, log):
    sim = Simulator(sd)
    systems = []
    for s in range(steps):
        st = sim.step()
        x, _ = sim.get_state()
        C = sim.constraint_set(x)
        systems.append((x.copy(), C.copy(), sim.scalars()["kappa"]))
    log("  %s: %d systems, last step %d contacts" % (sd.name, len(systems), st["contacts"]))
    return sim, systems


def train(steps=1500, lr=2e-3, lr_final=2e-4, n_probes=4, seed=0, sim_steps=12, log=print, n_stack=0, init=None):
    t0 = time.time()
    data = []
    for sd in training_scenes(seed, n_stack=n_stack):
        sim, systems = collect(sd, getattr(sd, "train_steps", sim_steps), log)
        data.append((sim, systems))
    log("collected %d systems from %d scenes in %.1f s" % (sum(len(s) for _, s in data), len(data), time.time() - t0))
    data[0][0].precond_init(seed + 1)
    if init is not None:
        data[0][0].precond_set_weights(init)
    state = data[0][0].precond_get_state()
    active = None
    losses = []
    rng = np.random.default_rng(seed)
    order = [(i, j) for i, (_, s) in enumerate(data) for j in range(len(s))]
    for it in range(steps):
        i, j = order[rng.integers(len(order))]
        sim, systems = data[i]
        if active is not i:
            if active is not None:
                state = data[active][0].precond_get_state()
            sim.precond_set_state(state)
            active = i
        x, C, kappa = systems[j]
        # cosine decay of the learning rate
        a = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * it / max(1, steps - 1)))
        losses.append(sim.precond_train_step(x, C, kappa, lr=a, n_probes=n_probes, seed=seed * 100003 + it))
        if (it + 1) % 100 == 0:
            log("step %d  loss %.4f (mean of last 100)  %.1f s" % (it + 1, float(np.mean(losses[-100:])), time.time() - t0))
    weights = data[active][0].precond_weights()
    return weights, np.array(losses), data


def save(path, weights, losses, meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, weights=weights, losses=losses.astype(np.float32), meta=np.array(repr(meta)))


def load(path=WEIGHTS):
    with np.load(path, allow_pickle=False) as f:
        return f["weights"].astype(np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--probes", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--sim-steps", type=int, default=12)
    ap.add_argument("--stack", type=int, default=0, help="four-sheet crops stepped into deep contact")
    ap.add_argument("--init", default=None, help="start from these weights (npz)")
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args()
    w, losses, _ = train(args.steps, args.lr, n_probes=args.probes, seed=args.seed, sim_steps=args.sim_steps,
                         log=lambda m: print(m, flush=True), n_stack=args.stack, init=load(args.init) if args.init else None)
    save(args.out, w, losses, dict(steps=args.steps, lr=args.lr, probes=args.probes, seed=args.seed, sim_steps=args.sim_steps,
                                   stack=args.stack, init=args.init))
    print("saved", args.out, flush=True)


if __name__ == "__main__":
    main()

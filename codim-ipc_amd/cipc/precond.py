"""Training of the learned preconditioner on the Newton systems the simulator produces.

Training data (--spec c4, the default): scenes of the two benchmark families -- the 41x41
single sheet over the icosphere and ground (2.5 cm spacing) and four-sheet 61x61 crops at
the 320k stack's 5 mm spacing -- dropped from seeded heights uniform in [0.2, 0.8] m with
Young's moduli log-uniform in [1e5, 1e7] Pa; a seeded fifth of the scenes also carries
lagged friction (mu 0.4) and another fifth the 1.01 strain limit.  Each scene is stepped
by the GPU simulator until it has lain in contact for `contact_steps` steps; every step's
end state, its active constraint set and kappa define one Newton system, until `systems`
systems are collected.  With --c2-steps K > 0 the 320k-triangle stack itself is added:
its 26 free-fall steps are run, then the end states of the next K steps (the sheets
landing on the sphere and on each other, 10^5 - 5 x 10^5 contacts) are systems too, and
a fraction --p-c2 of the training steps draws from them.  (--spec small: the earlier set
of nine near-contact scenes.)
The network is trained on all systems in rotation with Adam on the Kaporin loss
(cipc_precond_train_step, twice the probes on the large stack); weights and optimiser
state move between scenes through get/set state.  Training stops after --steps steps or
--budget-s seconds.  The result is saved as weights/precond.npz with the run's arguments
appended to the history of the weights it started from (--init).

Usage:  python -m cipc.precond [--spec c4|small] [--systems N] [--c2-steps K] [--p-c2 P]
                               [--steps N] [--budget-s S] [--init PATH] [--out PATH]
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


def c4_scenes(seed=0, n=60, frac_c2=1 / 3, frac_friction=0.2, frac_strain=0.2):
    """`n` scenes: 41x41 sheets and four-sheet 61x61 crops (5 mm spacing), drop height
    uniform in [0.2, 0.8] m (at least 2 mm above the sphere top), E log-uniform in
    [1e5, 1e7] Pa; seeded fractions with friction (mu 0.4) or the 1.01 strain limit"""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        y0 = float(max(rng.uniform(0.2, 0.8), 0.202))
        E = float(10 ** rng.uniform(5, 7))
        if rng.uniform() < frac_c2:
            sd = scenes.cloth_on_sphere(sheets=4, n=61, y0=y0, dy=0.005, seed=5000 + k, name="c4_s61_%d" % k)
            sd = _crop(sd, rng)
        else:
            sd = scenes.cloth_on_sphere(sheets=1, n=41, y0=y0, seed=4000 + k, name="c4_c41_%d" % k)
        sd.params["E"] = E
        u = rng.uniform()
        if u < frac_friction:
            sd.params.update(mu=0.4, epsv2=1e-6, fricIterAmt=1)
        elif u < frac_friction + frac_strain:
            sd.params.update(kappa_s=(1e3, 0.0), s=(1.01, 0.0), sHat=(1.0, 1.0))
        out.append(sd)
    return out


def collect_until_contact(sd, contact_steps, max_steps, log):
    """step until the scene has been in contact for `contact_steps` steps (or max_steps):
    one system per step"""
    sim = Simulator(sd, precision="fp32")   # the systems are states; the solve form only sets the cost
    systems = []
    in_contact = 0
    for s in range(max_steps):
        st = sim.step()
        x, _ = sim.get_state()
        C = sim.constraint_set(x)
        systems.append((x.copy(), C.copy(), sim.scalars()["kappa"]))
        in_contact += 1 if len(C) else 0
        if in_contact >= contact_steps:
            break
    log("  %s: %d systems (%d in contact), last step %d contacts" % (sd.name, len(systems), in_contact, st["contacts"]))
    return sim, systems


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


def collect_c2(steps, log, free_fall=26):
    """the 320k-triangle stack: `free_fall` untimed steps, then the end states of the next
    `steps` steps (FP32 block-Jacobi form; the systems are states, the form sets only the
    cost)"""
    sim = Simulator(scenes.scene_c2(), precision="fp32")
    for _ in range(free_fall):
        sim.step()
    systems = []
    for k in range(steps):
        st = sim.step()
        x, _ = sim.get_state()
        C = sim.constraint_set(x)
        systems.append((x.copy(), C.copy(), sim.scalars()["kappa"]))
        log("  c2 step %d: %d contacts, %d Newton iterations" % (free_fall + k + 1, len(C), st["newton_iters"]))
    return sim, systems


def train(steps=1500, lr=2e-3, lr_final=2e-4, n_probes=4, seed=0, sim_steps=12, log=print, n_stack=0, init=None, spec="c4",
          systems=5000, contact_steps=30, n_scenes=60, c2_steps=0, p_c2=0.5, budget_s=None):
    t0 = time.time()
    data = []
    if c2_steps > 0:
        data.append(collect_c2(c2_steps, log))
    if spec == "c4":
        total = 0
        for sd in c4_scenes(seed, n=n_scenes):
            if budget_s is not None and time.time() - t0 > 0.45 * budget_s:
                break
            sim, sys_ = collect_until_contact(sd, contact_steps, 90, log)
            data.append((sim, sys_))
            total += len(sys_)
            if total >= systems:
                break
    else:
        for sd in training_scenes(seed, n_stack=n_stack):
            sim, sys_ = collect(sd, getattr(sd, "train_steps", sim_steps), log)
            data.append((sim, sys_))
    log("collected %d systems from %d scenes in %.1f s" % (sum(len(s) for _, s in data), len(data), time.time() - t0))
    data[0][0].precond_init(seed + 1)
    if init is not None:
        data[0][0].precond_set_weights(init)
    state = data[0][0].precond_get_state()
    active = 0
    losses = []
    rng = np.random.default_rng(seed)
    big = 1 if c2_steps > 0 else 0          # data[0] is the stack when c2_steps > 0
    order = [(i, j) for i, (_, s) in enumerate(data) if i >= big for j in range(len(s))]
    done = 0
    for it in range(steps):
        if budget_s is not None and time.time() - t0 > budget_s:
            log("time budget reached after %d training steps" % it)
            break
        if big:
            # the stack with probability p_c2, else a uniformly drawn training scene; then a
            # uniformly drawn system of it
            i = 0 if (len(data) == 1 or rng.uniform() < p_c2) else 1 + int(rng.integers(len(data) - 1))
            j = int(rng.integers(len(data[i][1])))
        else:
            i, j = order[rng.integers(len(order))]
        sim, systems_i = data[i]
        if active != i:
            state = data[active][0].precond_get_state()
            sim.precond_set_state(state)
            active = i
        x, C, kappa = systems_i[j]
        # cosine decay of the learning rate
        a = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * it / max(1, steps - 1)))
        probes = n_probes * (2 if (big and i == 0) else 1)
        losses.append(sim.precond_train_step(x, C, kappa, lr=a, n_probes=probes, seed=seed * 100003 + it))
        done = it + 1
        if done % 100 == 0:
            log("step %d  loss %.4f (mean of last 100)  %.1f s" % (done, float(np.mean(losses[-100:])), time.time() - t0))
    weights = data[active][0].precond_weights()
    return weights, np.array(losses), data, done


def save(path, weights, losses, meta, history=()):
    """weights + the loss trace + meta: the list of training runs that produced them (the
    runs of the weights this one started from, then this run)"""
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, weights=weights, losses=losses.astype(np.float32), meta=np.array(repr(list(history) + [meta])))


def history(path):
    """the training runs recorded with saved weights (a list of dicts)"""
    import ast
    with np.load(path, allow_pickle=False) as f:
        m = ast.literal_eval(str(f["meta"]))
    return m if isinstance(m, list) else [m]


def load(path=WEIGHTS):
    with np.load(path, allow_pickle=False) as f:
        return f["weights"].astype(np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=2e-4)
    ap.add_argument("--probes", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--sim-steps", type=int, default=12)
    ap.add_argument("--stack", type=int, default=0, help="four-sheet crops stepped into deep contact (spec small)")
    ap.add_argument("--init", default=None, help="start from these weights (npz)")
    ap.add_argument("--spec", default="c4", choices=["c4", "small"])
    ap.add_argument("--systems", type=int, default=5000, help="Newton systems to collect (spec c4)")
    ap.add_argument("--scenes", type=int, default=60, help="c4 scenes at most (spec c4)")
    ap.add_argument("--contact-steps", type=int, default=30, help="steps in contact per scene (spec c4)")
    ap.add_argument("--c2-steps", type=int, default=0, help="contact steps of the 320k-triangle stack in the data")
    ap.add_argument("--p-c2", type=float, default=0.5, help="fraction of training steps on the stack's systems")
    ap.add_argument("--budget-s", type=float, default=None, help="wall-clock limit of collection + training")
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args()
    w, losses, data, done = train(args.steps, args.lr, lr_final=args.lr_final, n_probes=args.probes, seed=args.seed,
                                  sim_steps=args.sim_steps, log=lambda m: print(m, flush=True), n_stack=args.stack,
                                  init=load(args.init) if args.init else None, spec=args.spec, systems=args.systems,
                                  contact_steps=args.contact_steps, n_scenes=args.scenes, c2_steps=args.c2_steps, p_c2=args.p_c2,
                                  budget_s=args.budget_s)
    big = 1 if args.c2_steps > 0 else 0
    meta = dict(steps=done, lr=args.lr, lr_final=args.lr_final, probes=args.probes, seed=args.seed, spec=args.spec,
                c2_systems=len(data[0][1]) if big else 0, p_c2=args.p_c2 if big else 0.0,
                systems=sum(len(s) for _, s in data[big:]), scenes=len(data) - big, init=os.path.basename(args.init) if args.init else None)
    save(args.out, w, losses, meta, history(args.init) if args.init else ())
    print("saved", args.out, flush=True)


if __name__ == "__main__":
    main()

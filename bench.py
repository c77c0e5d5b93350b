"""Benchmark: implicit-Euler steps/s of the 320k-triangle cloth stack on one MI355X.

Workload (default): four 201x201 sheets 5 mm apart (161,604 cloth vertices / 320,000
triangles) falling onto a Dirichlet icosphere (r 0.2 m, level 4) and a 2-triangle ground,
rho 200, E 1e6, nu 0.3, t 1 mm, d_hat 1 mm, dt 0.01 s, Newton tol 1e-2 * bbox diagonal,
Newton linear solve to relative residual 1e-5.  A "step" is one call of the implicit-Euler
step (the reference's Advance_One_Step_IE_Hinge).  Inputs are generated on the host and
uploaded before the timed region; positions stay resident in HBM.  The default timed
window is steps 27-30 of the 50-step run: the 26 warm-up steps are the free fall (about
4 Newton iterations per step, no contact), the timed ones the sheets landing on the
sphere and on each other (tens of thousands of active constraints, about 10 Newton
iterations per step); the warm-up rate is reported beside the timed one.

Solver forms timed on the same scene, each from the initial state:
  fp64           block-Jacobi PCG in FP64;
  mixed          FP32 Hessian blocks, FP64 Krylov vectors / SpMV accumulation;
  fp32           FP32 Hessian blocks, Krylov vectors and SpMV accumulation;
  mixed_learned  the mixed form preconditioned by the learned factor (weights trained by
  fp32_learned   cipc.precond, codim-ipc_amd/cipc/weights/precond.npz), falling back to
                 block Jacobi when its iteration budget runs out; the same for fp32;
the reduced ones inside FP64 iterative refinement stopped on the FP64 residual (same
1e-5 relative bar).  `value` is the steps/s of the fastest reduced-precision form (named
in `mode`); every form's figures are reported under `forms`.

The JSON line carries the roofline of the dominant kernel of the headline form (the PCG
SpMV, timed from its own dispatch timestamps on the library's stream during the timed
steps) and a CPU baseline: the FP64 oracle (a C++ restatement of the reference path, same
PCG) on a bounded sample -- the first step of the same scene (free fall: the oracle would
need minutes per contact step) -- on the host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "codim-ipc_amd"))

METRIC = "sim steps/s + ms/Newton iter, 320k-tri cloth; plain FP64 vs mixed+learned PCG"
HBM_PEAK_GBS = 8000.0
KERNEL = {"fp64": ("k_pcg_spmv<double,double,%d,256,true,%d>", "PCG SpMV, FP64 blocks, FP64 vectors"),
          "mixed": ("k_pcg_spmv<float,double,%d,256,true,%d>", "PCG SpMV, FP32 blocks, FP64 vectors"),
          "fp32": ("k_pcg_spmv<float,float,%d,256,true,%d>", "PCG SpMV, FP32 blocks, FP32 vectors"),
          "mixed_learned": ("k_pcg_spmv<float,double,%d,256,true,%d>", "PCG SpMV, FP32 blocks, FP64 vectors"),
          "fp32_learned": ("k_pcg_spmv<float,float,%d,256,true,%d>", "PCG SpMV, FP32 blocks, FP32 vectors")}


def kernel_name(precision, tuning):
    """the PCG SpMV instance a run launches (template: blocks, vectors, lanes per block row,
    workgroup size, direction update, blocks per lane pass)"""
    name, what = KERNEL[precision]
    return "%s (%s)" % (name % (tuning["spmv_lanes"], tuning["spmv_unroll"]), what)
MODE = {"fp64": "plain FP64, block-Jacobi PCG",
        "mixed": "FP32 Hessian blocks, FP64 vectors, FP64 iterative refinement, block-Jacobi PCG",
        "fp32": "FP32 blocks/vectors/SpMV accumulation, FP64 iterative refinement, block-Jacobi PCG",
        "mixed_learned": "FP32 Hessian blocks, FP64 vectors, FP64 iterative refinement, learned-factor PCG (block-Jacobi fallback)",
        "fp32_learned": "FP32 blocks/vectors/SpMV accumulation, FP64 iterative refinement, learned-factor PCG (block-Jacobi fallback)"}
PRECISION = {"fp64": "fp64", "mixed": "mixed", "fp32": "fp32", "mixed_learned": "mixed", "fp32_learned": "fp32"}


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed FETCH_SIZE /
    WRITE_SIZE summary (profiles/), or None when no measurement matches."""
    path = os.path.join(ROOT, "profiles", "k_pcg_spmv_c2_hbm_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    if not kernel.startswith(t["kernel"] + " ") or t["workload"] != workload:
        return None, None
    return int(t["traffic_bytes_per_launch"]), t["summary"]


def gpu_run(args, sd, form):
    import cipc
    import numpy as np
    precision = form
    sim = cipc.Simulator(sd, pcg_tol=args.pcg_tol, precision=PRECISION[form])
    if form.endswith("_learned"):
        from cipc import precond
        sim.precond_set_weights(precond.load())
        sim.set_preconditioner("learned", budget=args.learned_budget, reuse=args.learned_reuse)
    if args.autotune:
        sim.autotune()          # settles during the warm-up steps
    tw = time.perf_counter()
    wstats = []
    for k in range(args.warmup):
        wstats.append(sim.step())
        print("[%s] warm-up step %d/%d" % (form, k + 1, args.warmup), file=sys.stderr, flush=True)
    tw = time.perf_counter() - tw
    sim.set_profiling(True)
    stats = []
    learned = form.endswith("_learned")
    info0 = sim.precond_info() if learned else None
    t0 = time.perf_counter()
    for k in range(args.steps):
        stats.append(sim.step())
        print("[%s] step %d/%d" % (form, k + 1, args.steps), file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    info1 = sim.precond_info() if learned else None
    prof = sim.profile()
    sim.set_profiling(False)
    x, _ = sim.get_state()
    C = sim.constraint_set(x)
    min_d = float(np.sqrt(sim.min_dist2(x, C)[1])) if len(C) else None
    newton = sum(s["newton_iters"] for s in stats)
    pcg = sum(s["pcg_iters"] for s in stats)
    stage_ms = {k[2:-3]: round(sum(s[k] for s in stats) / args.steps, 2)
                for k in ("t_total_ms", "t_constraint_set_ms", "t_ccd_ms", "t_hessian_ms", "t_solve_ms", "t_energy_ms", "t_gradient_ms")}
    roof = None
    if prof["launches"]:
        avg_s = prof["total_ms"] / prof["launches"] / 1e3
        ach = prof["bytes_per_launch"] / avg_s / 1e9
        kname = kernel_name(precision, sim.tuning())
        traffic, source = measured_traffic(kname, sd.name)
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4),
                    traffic=traffic, traffic_source=source, kernel=kname, launches_timed=prof["launches"],
                    pcg_iterations=pcg,
                    avg_us=round(avg_s * 1e6, 2),
                    bytes_per_launch=int(prof["bytes_per_launch"]), nnzb=prof["nnzb"], block_rows=prof["block_rows"])
    wn = sum(s["newton_iters"] for s in wstats[1:])
    return dict(precision=precision, steps_per_s=args.steps / elapsed,
                warmup_steps_per_s=(args.warmup / tw) if args.warmup else None, tuning=sim.tuning(),
                warmup_pcg_iters_per_newton=(sum(s["pcg_iters"] for s in wstats[1:]) / wn) if wn else None, ms_per_step=1e3 * elapsed / args.steps,
                ms_per_newton_iter=1e3 * elapsed / max(newton, 1), newton_iters_per_step=newton / args.steps,
                pcg_iters_per_newton=pcg / max(newton, 1), stage_ms_per_step=stage_ms, contacts_last_step=stats[-1]["contacts"],
                statuses=sorted(set(s["status"] for s in stats)), min_dist_final=min_d,
                learned_solves=(info1["solves"] - info0["solves"]) if learned else None,
                learned_fallbacks=(info1["fallbacks"] - info0["fallbacks"]) if learned else None,
                spmv_avg_us=roof["avg_us"] if roof else None, spmv_frac=roof["frac"] if roof else None, roofline=roof,
                _x=x, _energy=stats[-1]["energy"])


def batch_run(args):
    """C5: `--batch-count` independent scenes stepped together; aggregate steps/s and
    per-scene step latency percentiles over the timed batch steps."""
    import numpy as np
    import cipc
    from cipc import precond, scenes
    sds = scenes.scene_c5(count=args.batch_count)
    t0 = time.perf_counter()
    sims = []
    for sd in sds:
        sim = cipc.Simulator(sd, pcg_tol=args.pcg_tol, precision=PRECISION[args.batch_form])
        if args.batch_form.endswith("_learned"):
            sim.precond_set_weights(precond.load())
            sim.set_preconditioner("learned", budget=args.learned_budget)
        sims.append(sim)
    setup_s = time.perf_counter() - t0
    b = cipc.BatchStepper(sims, threads=args.batch_threads)
    for k in range(args.warmup):
        b.step()
        print("[batch] warm-up step %d/%d" % (k + 1, args.warmup), file=sys.stderr, flush=True)
    lat = []
    newton = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        st, ms = b.step()
        lat.append(ms)
        newton += sum(s["newton_iters"] for s in st)
        print("[batch] step %d/%d" % (k + 1, args.steps), file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    lat = np.concatenate(lat)
    n = len(sims)
    tris = [sd.nT for sd in sds]
    return {
        "metric": "aggregate sim steps/s, %d-scene batch (per-scene p50/p99 step latency)" % n,
        "value": round(n * args.steps / el, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / args.steps, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 + f64 refinement", "fp32_learned": "f32 + f64 refinement, bf16 network"}[args.batch_form],
        "data": "synthetic",
        "config": {"workload": "c5_batch", "scenes": n, "triangles_min": int(min(tris)), "triangles_max": int(max(tris)),
                   "triangles_total": int(sum(tris)), "threads": args.batch_threads, "form": args.batch_form},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2), "p99": round(float(np.percentile(lat, 99)), 2),
                       "max": round(float(lat.max()), 2)},
        "newton_iters_per_scene_step": round(newton / (n * args.steps), 2), "setup_s": round(setup_s, 1),
    }


def cpu_baseline(args, sd):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    orc = O.OracleScene(sd, solver=1, pcg_tol=args.pcg_tol, threads=threads)
    t0 = time.perf_counter()
    newton = 0
    for _ in range(args.cpu_steps):
        rc, st = orc.step()
        if rc <= 0:
            raise RuntimeError("oracle step failed: %d" % rc)
        newton += int(st["newton_iters"])
    el = time.perf_counter() - t0
    return dict(value=args.cpu_steps / el, unit="steps/s", cores=threads, kind="port",
                ms_per_newton_iter=1e3 * el / max(newton, 1), newton_iters_per_step=newton / args.cpu_steps,
                sample="first %d step(s) of the same scene on the FP64 C++ oracle (block-Jacobi PCG, rel-res %g), "
                       "%d OpenMP threads; %d Newton iterations in %.1f s" % (args.cpu_steps, args.pcg_tol, threads, newton, el))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # default window: steps 27-30 of the 50-step run, the sheets hitting the sphere and each
    # other (the first 26 steps are the free fall, run untimed)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=26)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c5"],
                    help="c2 (default): 320k-triangle stack; c1: 20k-triangle stack; c3: the c1 stack with "
                         "200,000 rod segments and 100,000 particles; c5: batch of independent scenes")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default launch shapes instead of tuning them during the warm-up")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--batch-threads", type=int, default=16)
    ap.add_argument("--batch-form", default="fp32_learned", choices=["fp64", "fp32", "fp32_learned"])
    ap.add_argument("--precision", default="all", choices=["fp64", "mixed", "fp32", "mixed_learned", "fp32_learned", "all"])
    ap.add_argument("--learned-budget", type=int, default=2500)
    ap.add_argument("--learned-reuse", type=int, default=None,
                    help="Newton solves per network evaluation (default 1: every solve)")
    ap.add_argument("--pcg-tol", type=float, default=1e-5)
    ap.add_argument("--mu", type=float, default=0.0,
                    help="friction coefficient (lagged friction, epsv2 = 1e-6, basis once per step); 0: frictionless")
    ap.add_argument("--strain-limit", type=float, default=0.0,
                    help="isotropic strain limit s (kappa_s = 1e3, sHat = 1); 0: off")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs one scene on one GPU; --gpus %d is not supported (use --gpus 1)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "c5":
        print(json.dumps(batch_run(args)))
        return
    from cipc import scenes
    sd = {"c1": scenes.scene_c1, "c2": scenes.scene_c2, "c3": scenes.scene_c3}[args.workload]()
    if args.mu > 0:
        sd.params.update(mu=args.mu, epsv2=1e-6, fricIterAmt=1)
    if args.strain_limit > 0:
        sd.params.update(kappa_s=(1e3, 0.0), s=(args.strain_limit, 0.0), sHat=(1.0, 1.0))

    modes = ["fp64", "mixed", "fp32"] if args.precision == "all" else [args.precision]
    if args.precision == "all":
        from cipc import precond
        if os.path.exists(precond.WEIGHTS):
            modes += ["mixed_learned", "fp32_learned"]
    runs = {m: gpu_run(args, sd, m) for m in modes}
    # each reduced-precision form against the FP64 form after the same steps from the same
    # state: max position difference relative to the FP64 displacement, and the last
    # step's incremental-potential difference relative to its magnitude
    if "fp64" in runs:
        import numpy as np
        x64, e64 = runs["fp64"]["_x"], runs["fp64"]["_energy"]
        disp = float(np.abs(x64 - sd.X.ravel()).max())
        for m, r in runs.items():
            r["vs_fp64_max_dx_rel"] = float(np.abs(r["_x"] - x64).max()) / max(disp, 1e-300)
            r["vs_fp64_energy_rel"] = abs(r["_energy"] - e64) / max(abs(e64), 1e-300)
    for r in runs.values():
        r.pop("_x")
        r.pop("_energy")
    reduced = [m for m in modes if m != "fp64"] or modes
    head = runs[max(reduced, key=lambda m: runs[m]["steps_per_s"])]
    cpu = None if args.no_cpu_baseline else cpu_baseline(args, sd)
    out = {
        "metric": METRIC,
        "value": round(head["steps_per_s"], 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "ms_per_newton_iter": round(head["ms_per_newton_iter"], 3),
        "newton_iters_per_step": round(head["newton_iters_per_step"], 3),
        "pcg_iters_per_newton": round(head["pcg_iters_per_newton"], 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "f32 blocks + f64", "fp32": "f32 + f64 refinement",
                  "mixed_learned": "f32 blocks + f64, bf16 network", "fp32_learned": "f32 + f64 refinement, bf16 network"}[head["precision"]],
        "data": "synthetic",
        "mode": MODE[head["precision"]],
        "config": {"workload": sd.name, "vertices": sd.nV, "triangles": sd.nT, "cloth_vertices": sd.n_cloth,
                   "rod_segments": 0 if sd.rods is None else len(sd.rods), "particles": 0 if sd.particles is None else len(sd.particles),
                   "steps_simulated_per_form": args.warmup + args.steps, "dt": sd.params["dt"], "dhat": sd.params["dHat"],
                   "pcg_tol": args.pcg_tol, "friction_mu": args.mu, "strain_limit": args.strain_limit},
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
        "forms": {m: {k: (round(v, 4) if isinstance(v, float) and k != "min_dist_final" and not k.startswith("vs_fp64") else v)
                      for k, v in r.items() if k != "roofline"}
                  for m, r in runs.items()},
    }
    if "fp64" in runs and len(runs) > 1:
        out["speedup_vs_fp64"] = round(head["steps_per_s"] / runs["fp64"]["steps_per_s"], 3)
        out["roofline_fp64"] = runs["fp64"]["roofline"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()

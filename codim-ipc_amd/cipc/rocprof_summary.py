"""Summaries of rocprofv3 databases (rocpd SQLite) as short markdown, and the copy kernel
used to calibrate the HBM byte counters.

  python -m cipc.rocprof_summary stats <run_results.db> [--top N]
      per-kernel calls / total / average duration (the top_kernels view, microseconds)
  python -m cipc.rocprof_summary pmc <run_results.db> <COUNTER> [--match SUBSTR] [--last N] [--min V]
      per-kernel mean of one counter (FETCH_SIZE / WRITE_SIZE are KiB per dispatch), over
      every dispatch, each kernel's last N, or the dispatches reading at least V
  python -m cipc.rocprof_summary tail <run_results.db> <SUBSTR> <N>
      mean / min / max duration (microseconds) of the last N dispatches of each kernel
      whose name contains SUBSTR (e.g. the launches of the timed steps after a warm-up)
  python -m cipc.rocprof_summary rates <trace_results.db> <pmc_results.db> [<pmc_results.db> ...] --match A,B,...
      per kernel whose name contains one of the substrings: mean duration (from the trace
      database) and the FP64 / FP32 VALU and BF16 MFMA rates the counter passes imply
      (64 x SQ_INSTS_VALU_FLOPS_FP64 (+ _TRANS), 64 x SQ_INSTS_VALU_FLOPS_FP32 (+ _TRANS),
      SQ_INSTS_VALU_MFMA_MOPS_BF16 x 512 FLOP, per dispatch, divided by the mean duration)
      against MI355X's peaks: 78.6 TFLOP/s FP64 vector, 157.3 TFLOP/s FP32 vector,
      2.5 PFLOP/s BF16 dense MFMA
  python -m cipc.rocprof_summary copy [--mib M] [--reps R]
      runs a plain device-to-device copy of M MiB R times on cuda:0 (torch's elementwise
      copy kernel): one read and one write of M MiB per launch, so FETCH_SIZE and
      WRITE_SIZE of these launches give the scale from counter KiB to HBM bytes.
"""
import argparse
import re
import sqlite3
import sys


def short_name(name):
    return re.sub(r"\(.*", "", name).replace("void ", "").strip()


def stats(db_path, top=25, out=sys.stdout):
    db = sqlite3.connect(db_path)
    rows = db.execute("select name,total_calls,total_duration,average,percentage from top_kernels").fetchall()
    tot = sum(r[2] for r in rows)
    print("| kernel | calls | total ms | avg us | % |", file=out)
    print("|---|---:|---:|---:|---:|", file=out)
    for name, calls, td, avg, pct in rows[:top]:
        print("| %s | %d | %.2f | %.2f | %.1f |" % (short_name(name), calls, td / 1e3, avg, pct), file=out)
    print("\nTotal kernel time %.1f ms over %d kernels (durations in us)" % (tot / 1e3, len(rows)), file=out)


def pmc(db_path, counter, match=None, last=0, out=sys.stdout, vmin=None):
    """per-kernel mean of one counter; with `last`, only each kernel's last `last`
    dispatches (e.g. the timed steps after a warm-up); with `vmin`, only dispatches whose
    value is at least vmin (PCG launches issued after convergence return at once and
    read next to nothing)"""
    db = sqlite3.connect(db_path)
    q = "select kernel_name, value from counters_collection where counter_name = ? order by dispatch_id"
    per = {}
    for name, v in db.execute(q, (counter,)):
        per.setdefault(name, []).append(v)
    print("| kernel | dispatches | mean %s | min | max |" % counter, file=out)
    print("|---|---:|---:|---:|---:|", file=out)
    for name, vals in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        if match and match not in name:
            continue
        if last:
            vals = vals[-last:]
        if vmin is not None:
            vals = [v for v in vals if v >= vmin]
            if not vals:
                continue
        print("| %s | %d | %.1f | %.1f | %.1f |" % (short_name(name), len(vals), sum(vals) / len(vals), min(vals), max(vals)), file=out)


def tail(db_path, match, last, out=sys.stdout):
    db = sqlite3.connect(db_path)
    cols = [r[1] for r in db.execute("pragma table_info(kernels)")]
    name = "kernel_name" if "kernel_name" in cols else "name"
    per = {}
    q = "select %s, start, end from kernels order by start" % name
    for n, t0, t1 in db.execute(q):
        if match in n:
            per.setdefault(n, []).append((t1 - t0) / 1e3)
    print("| kernel | dispatches | mean us | min | max |", file=out)
    print("|---|---:|---:|---:|---:|", file=out)
    for n, d in per.items():
        d = d[-last:]
        print("| %s | %d | %.2f | %.2f | %.2f |" % (short_name(n), len(d), sum(d) / len(d), min(d), max(d)), file=out)


PEAK_GFLOPS = {"fp64": 78.6e3, "fp32": 157.3e3, "bf16_mfma": 2.5e6}


def _durations(db_path):
    db = sqlite3.connect(db_path)
    cols = [r[1] for r in db.execute("pragma table_info(kernels)")]
    name = "kernel_name" if "kernel_name" in cols else "name"
    per = {}
    for n, t0, t1 in db.execute("select %s, start, end from kernels" % name):
        per.setdefault(short_name(n), []).append((t1 - t0) / 1e3)
    return per


def _counters(db_paths):
    per = {}
    for p in db_paths:
        db = sqlite3.connect(p)
        for n, c, v in db.execute("select kernel_name, counter_name, value from counters_collection"):
            per.setdefault(short_name(n), {}).setdefault(c, []).append(v)
    return per


def rates(trace_db, pmc_dbs, matches, out=sys.stdout):
    dur = _durations(trace_db)
    cnt = _counters(pmc_dbs)
    mean = lambda v: sum(v) / len(v) if v else 0.0  # noqa: E731
    print("| kernel | dispatches | mean us | FP64 GFLOP/s (frac of 78.6 TF) | FP32 GFLOP/s (frac of 157.3 TF) | BF16 MFMA GFLOP/s (frac of 2.5 PF) | VALU insts / dispatch |",
          file=out)
    print("|---|---:|---:|---:|---:|---:|---:|", file=out)
    for k in sorted(cnt):
        if not any(m in k for m in matches):
            continue
        c = cnt[k]
        d = dur.get(k, [])
        us = mean(d)
        if us <= 0:
            continue
        # the VALU FLOPS counters count per wave instruction (FMA 2, others 1); x 64 lanes, as
        # the counter set's own TOTAL_64_OPS does with the instruction counts (full waves)
        f64 = 64.0 * (mean(c.get("SQ_INSTS_VALU_FLOPS_FP64", [])) + mean(c.get("SQ_INSTS_VALU_FLOPS_FP64_TRANS", [])))
        f32 = 64.0 * (mean(c.get("SQ_INSTS_VALU_FLOPS_FP32", [])) + mean(c.get("SQ_INSTS_VALU_FLOPS_FP32_TRANS", [])))
        bf = 512.0 * mean(c.get("SQ_INSTS_VALU_MFMA_MOPS_BF16", []))
        g = lambda fl: fl / (us * 1e-6) / 1e9  # noqa: E731
        print("| %s | %d | %.1f | %.0f (%.3f) | %.0f (%.3f) | %.0f (%.4f) | %.3g |" % (
            k, len(d), us, g(f64), g(f64) / PEAK_GFLOPS["fp64"], g(f32), g(f32) / PEAK_GFLOPS["fp32"], g(bf),
            g(bf) / PEAK_GFLOPS["bf16_mfma"], mean(c.get("SQ_INSTS_VALU", []))), file=out)


def copy(mib=256, reps=3):
    import torch
    n = (mib << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied %d MiB x %d" % (mib, reps), flush=True)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats"); s.add_argument("db"); s.add_argument("--top", type=int, default=25)
    p = sub.add_parser("pmc"); p.add_argument("db"); p.add_argument("counter"); p.add_argument("--match")
    p.add_argument("--last", type=int, default=0)
    p.add_argument("--min", type=float, default=None, dest="vmin")
    t = sub.add_parser("tail"); t.add_argument("db"); t.add_argument("match"); t.add_argument("last", type=int)
    c = sub.add_parser("copy"); c.add_argument("--mib", type=int, default=256); c.add_argument("--reps", type=int, default=3)
    r = sub.add_parser("rates"); r.add_argument("trace_db"); r.add_argument("pmc_dbs", nargs="+"); r.add_argument("--match", required=True)
    a = ap.parse_args()
    if a.cmd == "stats":
        stats(a.db, a.top)
    elif a.cmd == "pmc":
        pmc(a.db, a.counter, a.match, a.last, vmin=a.vmin)
    elif a.cmd == "tail":
        tail(a.db, a.match, a.last)
    elif a.cmd == "rates":
        rates(a.trace_db, a.pmc_dbs, a.match.split(","))
    else:
        copy(a.mib, a.reps)


if __name__ == "__main__":
    main()

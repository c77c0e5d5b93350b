"""Summaries of rocprofv3 databases (rocpd SQLite) as short markdown, and the copy kernel
used to calibrate the HBM byte counters.

  python -m cipc.rocprof_summary stats <run_results.db> [--top N]
      per-kernel calls / total / average duration (the top_kernels view, microseconds)
  python -m cipc.rocprof_summary pmc <run_results.db> <COUNTER> [--match SUBSTR] [--last N]
      per-kernel mean of one counter (FETCH_SIZE / WRITE_SIZE are KiB per dispatch), over
      every dispatch or each kernel's last N
  python -m cipc.rocprof_summary tail <run_results.db> <SUBSTR> <N>
      mean / min / max duration (microseconds) of the last N dispatches of each kernel
      whose name contains SUBSTR (e.g. the launches of the timed steps after a warm-up)
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


def pmc(db_path, counter, match=None, last=0, out=sys.stdout):
    """per-kernel mean of one counter; with `last`, only each kernel's last `last`
    dispatches (e.g. the timed steps after a warm-up)"""
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
    t = sub.add_parser("tail"); t.add_argument("db"); t.add_argument("match"); t.add_argument("last", type=int)
    c = sub.add_parser("copy"); c.add_argument("--mib", type=int, default=256); c.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    if a.cmd == "stats":
        stats(a.db, a.top)
    elif a.cmd == "pmc":
        pmc(a.db, a.counter, a.match, a.last)
    elif a.cmd == "tail":
        tail(a.db, a.match, a.last)
    else:
        copy(a.mib, a.reps)


if __name__ == "__main__":
    main()

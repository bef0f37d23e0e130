"""Summaries of rocprofv3 output databases (kernel trace and hardware counters).

python -m cloverleaf_amd.profiling stats <results.db>
    per-kernel dispatch count and duration statistics (markdown table)
python -m cloverleaf_amd.profiling sq <run.db> [<run.db> ...]
    per-kernel shader-sequencer counters (SQ_*) averaged per dispatch, merged over
    several --pmc runs: the share of wave time waiting on memory / LDS counters
    (SQ_WAIT_ANY), waiting to issue (SQ_WAIT_INST_ANY) and issuing (SQ_ACTIVE_INST_*),
    and instructions per wave (SQ_INSTS_* / SQ_WAVES)
python -m cloverleaf_amd.profiling tcc <run.db>
    per-kernel L2 hits and misses (TCC_HIT_sum, TCC_MISS_sum) averaged per dispatch
python -m cloverleaf_amd.profiling traffic <fetch.db> <write.db> --probe-mib N --out file.json --form F
    calibrated HBM bytes per dispatch of every kernel, from one FETCH_SIZE and one
    WRITE_SIZE run of the same command; the counters of the copy probe kernel (an
    elementwise copy of N MiB launched first by `bench.py --copy-probe-mib N`) give the
    bytes-per-counter-unit factors.
"""
from __future__ import annotations

import argparse
import json
import sqlite3
import sys
from collections import defaultdict


def kernel_stats(db_path: str, split_at: str = "", labels=()):
    """Per-kernel duration statistics.  With `split_at`, the dispatches are cut into
    consecutive segments, each starting at a dispatch of a kernel whose name contains
    `split_at` (bench.py initialises a fresh table once per form, so k_init_table
    separates the forms; a segment ends where the next one's split kernel runs, so set-up
    copies and fills just before it count in the previous segment); the result is a list of (label, rows) per segment."""
    db = sqlite3.connect(db_path)
    disp = db.execute("select name, start, end from kernels order by start").fetchall()
    segs = [[]]
    for name, t0, t1 in disp:
        if split_at and split_at in name and any(split_at in n for n, _ in segs[-1]):
            segs.append([])
        segs[-1].append((name, t1 - t0))
    out = []
    for i, seg in enumerate(segs if split_at else [sum(segs, [])]):
        acc = defaultdict(list)
        for name, dt in seg:
            acc[name].append(dt)
        rows = [dict(name=k, count=len(v), avg_us=sum(v) / len(v) / 1e3, min_us=min(v) / 1e3, max_us=max(v) / 1e3,
                     total_us=sum(v) / 1e3) for k, v in acc.items()]
        rows.sort(key=lambda r: -r["total_us"])
        out.append((labels[i] if i < len(labels) else f"segment {i}", rows))
    return out if split_at else out[0][1]


def counter_per_kernel(db_path: str, counter: str):
    """(average, count, max) of the counter per dispatch, per kernel name (summed over the
    counter's instances)."""
    db = sqlite3.connect(db_path)
    rows = db.execute("select dispatch_id, kernel_name, sum(value) from counters_collection where counter_name = ? "
                      "group by dispatch_id", (counter,)).fetchall()
    acc = defaultdict(list)
    for _, name, v in rows:
        acc[name].append(v)
    return {k: (sum(v) / len(v), len(v), max(v)) for k, v in acc.items()}


def short_name(name: str) -> str:
    base = name.split("(")[0]
    return base.replace("void ", "").replace("ep::", "").strip()


def traffic(fetch_db: str, write_db: str, probe_mib: int):
    fetch = counter_per_kernel(fetch_db, "FETCH_SIZE")
    write = counter_per_kernel(write_db, "WRITE_SIZE")
    # the probe: the elementwise kernel that reads and writes probe_mib MiB (largest
    # single-dispatch FETCH among elementwise kernels)
    probe = [k for k in fetch if "elementwise" in k.lower()]
    if not probe:
        raise SystemExit("no elementwise probe kernel in the FETCH_SIZE run")
    pk = max(probe, key=lambda k: fetch[k][2])
    true_bytes = probe_mib * (1 << 20)
    f_fac = true_bytes / (fetch[pk][2] * 1024.0)  # counters are in KiB
    w_fac = true_bytes / (write[pk][2] * 1024.0) if pk in write and write[pk][2] > 0 else 1.0
    out = {}
    for k in fetch:
        if k == pk:
            continue
        fb = fetch[k][0] * 1024.0 * f_fac
        wb = write.get(k, (0.0, 0))[0] * 1024.0 * w_fac
        out[short_name(k)] = dict(fetch_bytes=round(fb), write_bytes=round(wb), dispatches=fetch[k][1])
    return out, dict(probe_kernel=short_name(pk), fetch_factor=f_fac, write_factor=w_fac, probe_mib=probe_mib)


def sq_summary(dbs):
    """Per kernel: SQ counters averaged per dispatch, merged over the given runs."""
    acc = defaultdict(lambda: defaultdict(list))
    for path in dbs:
        db = sqlite3.connect(path)
        rows = db.execute("select dispatch_id, kernel_name, counter_name, sum(value) from counters_collection "
                          "where counter_name like 'SQ_%' group by dispatch_id, counter_name").fetchall()
        for _, kn, cn, v in rows:
            acc[short_name(kn)][cn].append(v)
    out = []
    for k, cs in acc.items():
        d = {cn: sum(v) / len(v) for cn, v in cs.items()}
        out.append((k, d))
    out.sort(key=lambda kd: -(kd[1].get("SQ_WAIT_ANY", 0) + kd[1].get("SQ_ACTIVE_INST_ANY", 0)))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s1 = sub.add_parser("stats")
    s1.add_argument("db")
    s1.add_argument("--top", type=int, default=25)
    s1.add_argument("--split-at", default="", help="start a new table at each dispatch of this kernel")
    s1.add_argument("--labels", default="", help="comma-separated titles of the split tables")
    s4 = sub.add_parser("tcc")
    s4.add_argument("db")
    s4.add_argument("--top", type=int, default=16)
    s3 = sub.add_parser("sq")
    s3.add_argument("dbs", nargs="+")
    s3.add_argument("--top", type=int, default=12)
    s2 = sub.add_parser("traffic")
    s2.add_argument("fetch_db")
    s2.add_argument("write_db")
    s2.add_argument("--probe-mib", type=int, required=True)
    s2.add_argument("--form", required=True)
    s2.add_argument("--out", required=True)
    s2.add_argument("--command", default="")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        tables = (kernel_stats(a.db, a.split_at, [x for x in a.labels.split(",") if x]) if a.split_at
                  else [("", kernel_stats(a.db))])
        for label, rows in tables:
            if label:
                print(f"\n### {label}\n")
            print("| kernel | dispatches | avg us | min us | max us | total us |")
            print("|---|---|---|---|---|---|")
            for r in rows[:a.top]:
                print(f"| `{short_name(r['name'])}` | {r['count']} | {r['avg_us']:.2f} | {r['min_us']:.2f} | "
                      f"{r['max_us']:.2f} | {r['total_us']:.0f} |")
        return 0
    if a.cmd == "tcc":
        hits = counter_per_kernel(a.db, "TCC_HIT_sum")
        miss = counter_per_kernel(a.db, "TCC_MISS_sum")
        rows = []
        for k in hits:
            h, m = hits[k][0], miss.get(k, (0.0, 0, 0))[0]
            if h + m > 0:
                rows.append((short_name(k), hits[k][1], h, m))
        rows.sort(key=lambda r: -r[3])
        print("| kernel | dispatches | L2 hits (M requests) | L2 misses (M requests) | hit rate |")
        print("|---|---:|---:|---:|---:|")
        for k, n, h, m in rows[:a.top]:
            print(f"| `{k}` | {n} | {h / 1e6:.1f} | {m / 1e6:.1f} | {100 * h / (h + m):.1f}% |")
        return 0
    if a.cmd == "sq":
        print("| kernel | waves | wait on memory/LDS | wait to issue | issuing | VALU | LDS | VALU / VMEM rd / VMEM wr / LDS instr per wave |")
        print("|---|---:|---:|---:|---:|---:|---:|---|")
        for k, d in sq_summary(a.dbs)[:a.top]:
            t = d.get("SQ_WAIT_ANY", 0) + d.get("SQ_WAIT_INST_ANY", 0) + d.get("SQ_ACTIVE_INST_ANY", 0)
            w = d.get("SQ_WAVES", 0)
            if not t or not w:
                continue
            pct = lambda x: f"{100 * d.get(x, 0) / t:.0f}%"
            per = " / ".join(f"{d.get(x, 0) / w:.0f}" for x in ("SQ_INSTS_VALU", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR",
                                                               "SQ_INSTS_LDS"))
            print(f"| `{k}` | {w:.0f} | {pct('SQ_WAIT_ANY')} | {pct('SQ_WAIT_INST_ANY')} | {pct('SQ_ACTIVE_INST_ANY')} | "
                  f"{pct('SQ_ACTIVE_INST_VALU')} | {pct('SQ_ACTIVE_INST_LDS')} | {per} |")
        return 0
    ker, cal = traffic(a.fetch_db, a.write_db, a.probe_mib)
    try:
        with open(a.out) as fh:
            data = json.load(fh)
    except FileNotFoundError:
        data = {"kernels": {}}
    data.setdefault("kernels", {})
    data["calibration"] = cal
    if a.command:
        data.setdefault("commands", {})[a.form] = a.command
    best = {}  # base name -> the template instance with the most dispatches (the tuned one)
    for k, v in ker.items():
        data["kernels"][f"{a.form}:{k}"] = v
        b = k.split("<")[0]
        if b not in best or v["dispatches"] > best[b]["dispatches"]:
            best[b] = v
    for b, v in best.items():
        data["kernels"][f"{a.form}:{b}"] = v
    with open(a.out, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    for k, v in sorted(ker.items(), key=lambda kv: -(kv[1]["fetch_bytes"] + kv[1]["write_bytes"]))[:12]:
        print(f"{k}: fetch {v['fetch_bytes'] / 1e6:.1f} MB, write {v['write_bytes'] / 1e6:.1f} MB per dispatch "
              f"({v['dispatches']} dispatches)")
    return 0


if __name__ == "__main__":
    sys.exit(main())

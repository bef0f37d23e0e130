"""EP training-step benchmark on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ep-cfg4-50m] [--forms ...]

Workload (default): the 50M-node mixed-precision configuration: undirected Chung-Lu graph,
50M nodes, 1B edges (1.99B CSR entries), beta 2.2; features U{1..64} Zipf(1.1) over a 20M
vocabulary; d=256, B=16384, K=20 negatives, S=20 neighbours, margin 0.5, lr 0.01.
Synthetic, seeded inputs; random-init table.

Forms timed on the same graph:
  fp32-plain           FP32 table, one batch per step, one launch per kernel (no graph),
                       the default segment tile (no tuning), sort + segment-reduce scatter
  bf16-master-batched  BF16 table + persistent FP32 master rows, stochastic-rounding
                       write-back; --concurrent batches per step (they read the same table
                       and update in batch order, like the reference's concurrently running
                       batches), steps replayed from a captured hipGraph, tuned segment tile,
                       FP16 adjoint rows (--fp16-adjoints)
  bf16-batched         the same without the persistent master (each touched row is taken to
                       FP32 for its update)
  fp32-batched         FP32 table, batched as above
The JSON `value` is the examples/s of the headline form (bf16-master-batched, the
configuration's BF16 table + FP32 master + SR).  Every example is one anchor taken through
h(v), ~h(v), K negatives, the margin loss, the hand-written backward and its batch's
sparse Adam update.  A "step" in `ms_per_step` is one launch sequence of the headline form,
i.e. --concurrent reference batches.

`roofline` is the HBM roofline of the dominant kernel of the headline form (the
short-row update, or the forward gather when that takes longer): its algorithmic bytes
per launch over its average launch duration, timed with HIP events on the context's
stream.  `roofline_gather` is the forward's: the distinct feature rows one launch reads
(counted on the host from a step's sampled lists) over its event-timed duration.
`roofline_step` applies the algorithmic-byte model (rows with multiplicity) to the whole
step.  `cpu_baseline` times the C++ restatement of the reference's rayon loop on the host
cores (bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 (matrix; packed vector math reaches the same)
PHASES = ["begin+sample+scan", "forward", "emit", "sort", "plan", "short_update", "pieces+medium", "long_rows"]


def gather_bytes(rows, nodes, examples, dims, bytes_per_value):
    """Sum over examples of R_e*d*b + 4 R_e + 4 (S_e + K) + 16 (1 + S_e + K)."""
    return rows * dims * bytes_per_value + 4 * rows + 4 * nodes + 16 * (examples + nodes)


def update_bytes(urows, dims, bytes_per_value, master=False):
    """Per updated value: theta read + write (2b), m and v read + write (16); with a
    persistent FP32 master the master is read and written (8) and the table row only
    written (b)."""
    return urows * dims * ((bytes_per_value + 24) if master else (2 * bytes_per_value + 16))


def attention_flops(att_items_sq, att_items, heads, dk, dm):
    """Attention FLOP model over the timed steps: per block of n items and per head,
    forward 2 n^2 d_k (scores) + n^2 (softmax) + 2 n d_model (weighted values); the
    backward about twice the forward.  Returns (forward, forward + backward)."""
    fwd = heads * (2 * dk * att_items_sq + att_items_sq + 2 * dm * att_items)
    return fwd, 3 * fwd


def run_form(w, form, steps, warmup, scatter=None, concurrent=1, graph=True, fp16_adjoints=False):
    from cloverleaf_amd import _native as N

    cfg = N.default_config()
    cfg.dims, cfg.batch_size, cfg.negatives, cfg.gamma, cfg.alpha = w.dims, w.batch, w.negatives, w.gamma, w.alpha
    if w.att_dk:
        cfg.attention_dk, cfg.attention_heads = w.att_dk, w.att_heads
    cfg.max_nodes_kind, cfg.max_nodes = N.EP_SAMPLE_FIXED, w.max_nodes
    cfg.valid_pct = 0.0
    cfg.passes = max(1, w.passes)
    cfg.scatter = N.EP_SCATTER_ATOMIC if scatter == "atomic" else N.EP_SCATTER_SORT
    if form.endswith("-plain"):
        cfg.use_graph, cfg.concurrent_batches = 0, 1
        # default tile, no tuning: 4 values per lane, or 16 where a row needs more than 64 lanes
        cfg.seg_vpl, cfg.seg_unroll = (4, 8) if N.row_width(cfg) <= 256 else (16, 2)
    else:
        cfg.use_graph, cfg.concurrent_batches = (1 if graph else 0), concurrent
    prec = form.rsplit("-", 1)[0]
    cfg.precision = {"fp32": N.EP_PREC_FP32, "bf16": N.EP_PREC_BF16, "bf16-master": N.EP_PREC_BF16_MASTER}[prec]
    cfg.fp16_adjoints = 1 if (fp16_adjoints and prec != "fp32" and not w.att_dk and cfg.scatter == N.EP_SCATTER_SORT) else 0
    with N.Context(cfg) as ctx:
        ctx.set_graph(w.indptr, w.indices)
        ctx.set_features(w.f_indptr, w.f_ids, w.vocab)
        ctx.set_table(None)
        train = np.arange(w.n_nodes, dtype=np.uint32)
        ctx.set_split(train, np.zeros(0, dtype=np.uint32))
        r = ctx.bench_steps(warmup, steps, instrument=True)
        audit = gather_audit(ctx, w, int(cfg.concurrent_batches) * w.batch)
    b = 2 if prec.startswith("bf16") else 4
    S = max(r.steps, 1)
    ph = [float(x) for x in r.ms_phase]
    W = N.row_width(cfg)  # table row width (attention rows hold Q, K and V of every head)
    g_launch = gather_bytes(r.rows, r.nodes, r.examples, W, b) / S
    u_launch = update_bytes(r.short_rows, W, b, prec == "bf16-master") / S
    step_bytes = (gather_bytes(r.rows, r.nodes, r.examples, W, b)
                  + update_bytes(r.unique_rows, W, b, prec == "bf16-master")) / S
    extra = {}
    if w.att_dk:
        _, fb = attention_flops(r.att_items_sq, r.att_items, w.att_heads, w.att_dk, w.dims)
        extra = {"attention_blocks_items_sq_per_step": r.att_items_sq / S,
                 "attention_flops_per_step": fb / S,
                 "attention_tflops": fb / S / (ph[1] / 1e3) / 1e12 if ph[1] > 0 else 0.0}
    u_rows, occ, n_nodes = audit
    # the forward's distinct rows per launch (P batches), ids and node offsets
    g_unique = u_rows * W * b + 4 * occ + 16 * n_nodes
    return {
        "form": form,
        "gather_unique_rows_per_launch": int(u_rows),
        "gather_occurrences_per_launch": int(occ),
        "gather_unique_bytes_per_launch": g_unique,
        "gather_unique_gbs": g_unique / (ph[1] / 1e3) / 1e9 if ph[1] > 0 else 0.0,
        "scatter": scatter or "sort",
        "concurrent_batches": int(cfg.concurrent_batches),
        "adjoint_rows": "fp16" if cfg.fp16_adjoints else "fp32",
        "seg_tile": [int(r.seg_vpl), int(r.seg_unroll)],
        "forward_rows_in_flight_x": int(r.fwd_x),
        "examples_per_s": r.examples / (r.ms_total / 1e3),
        "ms_per_step": r.ms_total / S,
        "examples": int(r.examples),
        "rows_per_example": r.rows / max(r.examples, 1),
        "unique_rows_per_step": r.unique_rows / S,
        "short_rows_per_step": r.short_rows / S,
        "long_rows_last_step": int(r.last_long_rows),
        "pieces_last_step": int(r.last_pieces),
        "phase_ms": {k: round(v, 4) for k, v in zip(PHASES, ph)},
        "forward_bytes_per_launch": g_launch,
        "forward_gbs": g_launch / (ph[1] / 1e3) / 1e9 if ph[1] > 0 else 0.0,
        "update_bytes_per_launch": u_launch,
        "update_gbs": u_launch / (ph[5] / 1e3) / 1e9 if ph[5] > 0 else 0.0,
        "step_bytes": step_bytes,
        "step_gbs": step_bytes / (r.ms_total / S / 1e3) / 1e9,
        **extra,
    }


def gather_audit(ctx, w, n_anchors):
    """Distinct feature rows and feature occurrences the forward reads for one step of
    n_anchors anchors: the GPU sampler draws the step's neighbour and negative lists, the
    host expands every node of every example with a neighbourhood to its feature ids."""
    rng = np.random.default_rng(123)
    anchors = rng.choice(w.n_nodes, size=min(n_anchors, w.n_nodes), replace=False).astype(np.uint32)
    cnt, nbr, neg = ctx.sample_batch(anchors, 1)
    live = cnt > 0
    S = nbr.shape[1]
    picked = nbr[live][np.arange(S)[None, :] < cnt[live][:, None]]
    nodes = np.concatenate([anchors[live], picked, neg[live].ravel()]).astype(np.int64)
    fi = np.asarray(w.f_indptr, dtype=np.int64)
    starts, lens = fi[nodes], fi[nodes + 1] - fi[nodes]
    tot = int(lens.sum())
    idx = np.repeat(starts - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens) + np.arange(tot)
    return len(np.unique(np.asarray(w.f_ids)[idx])), tot, len(nodes)


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(w, budget_s=15.0):
    """Times the C++ restatement of the reference's rayon loop (oracle/_build/libep_cpu.so,
    ep_cpu_bench) on the host cores, in process on the workload's own arrays: examples of a
    batch in parallel, O(deg) reservoir neighbours, first-leaf hash maps, per-batch merge,
    sparse Adam, FP32.  The clock starts after the table initialisation; the sample is the
    batches finished within budget_s seconds (the last one may run past it)."""
    import ctypes
    from ctypes import POINTER, c_double, c_float, c_uint32, c_uint64

    path = os.path.join(ROOT, "oracle", "_build", "libep_cpu.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    u64p, u32p = POINTER(c_uint64), POINTER(c_uint32)
    lib.ep_cpu_bench.argtypes = [u64p, u32p, u64p, u32p, c_uint64, c_uint32, c_uint32, c_uint32, c_uint32, c_uint32,
                                 c_float, c_float, c_uint32, c_uint32, c_uint32, c_float, POINTER(c_double)]
    lib.ep_cpu_bench.restype = ctypes.c_int
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    arrs = [np.ascontiguousarray(w.indptr, dtype=np.uint64), np.ascontiguousarray(w.indices, dtype=np.uint32),
            np.ascontiguousarray(w.f_indptr, dtype=np.uint64), np.ascontiguousarray(w.f_ids, dtype=np.uint32)]
    ptrs = [a.ctypes.data_as(u64p if a.dtype == np.uint64 else u32p) for a in arrs]
    out = (c_double * 3)()
    t0 = time.time()
    rc = lib.ep_cpu_bench(*ptrs, w.n_nodes, w.vocab, w.dims, w.batch, w.negatives, w.max_nodes, w.gamma, w.alpha,
                          w.att_heads if w.att_dk else 0, w.att_dk, cores, budget_s, out)
    if rc != 0:
        return None
    examples, seconds, batches = out[0], out[1], int(out[2])
    print(f"[bench] cpu baseline: {examples:.0f} anchors in {seconds:.1f} s on {cores} threads "
          f"({time.time() - t0:.1f} s with the table initialisation)", file=sys.stderr, flush=True)
    return {"value": round(examples / seconds, 1), "unit": "examples/s", "cores": cores, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"{examples:.0f} anchors ({batches} batches of {w.batch}, each batch's examples, merge and "
                      f"sparse Adam over all {cores} threads) of the same workload, FP32, C++ restatement of the "
                      f"reference's rayon loop"
                      + (" (attention aggregator)" if w.att_dk else "")}


def pmc_traffic(kernel, workload, form):
    """Calibrated HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/hbm_traffic_<workload>.json, written from two rocprofv3 --pmc runs)."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        data = json.load(fh)
    ent = data.get("kernels", {}).get(f"{form}:{kernel}")
    if not ent:
        return None
    return ent["fetch_bytes"] + ent["write_bytes"]


def bench_align(w, args):
    """NeighborhoodAligner.align over every node of the workload's graph (edge weights:
    uniform rows), random FP32 node embeddings of --align-dims; the graph and embeddings
    are resident in HBM, each run aligns all nodes (kernel timed with HIP events)."""
    from cloverleaf_amd import _native as N
    from cloverleaf_amd.api import Graph

    g = Graph.from_csr(w.indptr, w.indices)
    cdf = g.edge_cdf()
    d = args.align_dims
    emb = np.random.default_rng(1).standard_normal((w.n_nodes, d), dtype=np.float32)
    deg = np.diff(w.indptr.astype(np.int64))
    m = np.minimum(deg, args.align_max_neighbors) if args.align_max_neighbors else deg
    algo = float(((m + 2) * d * 4 + 8 * m).sum())  # neighbour rows + own row + output row, edge id + probability
    ms = []
    with N.Aligner(w.indptr, w.indices, cdf, emb) as al:
        for i in range(args.warmup + args.steps):
            al.run(0.5, args.align_max_neighbors or None)
            if i >= args.warmup:
                ms.append(al.last_kernel_ms())
    avg = float(np.mean(ms))
    gbs = algo / (avg * 1e-3) / 1e9
    print(f"[bench] align: {w.n_nodes} nodes, {int(m.sum())} edges read, d={d}: {avg:.3f} ms per run, "
          f"{gbs:.0f} GB/s algorithmic", file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": "NeighborhoodAligner.align nodes/sec", "value": round(w.n_nodes / (avg * 1e-3), 1),
        "unit": "nodes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(avg, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded graph, random node embeddings)",
        "config": {"workload": w.name, "nodes": w.n_nodes, "csr_entries": int(len(w.indices)), "dims": d,
                   "max_neighbors": args.align_max_neighbors or None, "alpha": 0.5},
        "roofline": {"bound": "hbm", "kernel": "k_align", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": algo},
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="ep-cfg4-50m")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--graph-scale", type=float, default=1.0,
                    help="shrink only the graph (nodes, edges); vocabulary and model shape unchanged")
    ap.add_argument("--forms", default="fp32-plain,bf16-master-batched,bf16-batched")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--scatters", default="sort", help="scatter paths tried per form (sort,atomic; atomic runs one batch per step)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--concurrent", type=int, default=0,
                    help="reference batches per step in the batched forms (the plain forms use 1); default 8, "
                         "2 for attention workloads (one gradient row per feature occurrence)")
    ap.add_argument("--fp16-adjoints", type=int, default=1,
                    help="1: the BF16 forms hold the adjoint rows the update gathers in FP16 (per-row scale)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the batched forms eagerly (for hardware-counter runs)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (HBM-counter calibration)")
    ap.add_argument("--align", action="store_true",
                    help="measure NeighborhoodAligner.align (k_align) on the workload's graph instead of training")
    ap.add_argument("--align-dims", type=int, default=128)
    ap.add_argument("--align-max-neighbors", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is described, not built)", file=sys.stderr)
        sys.exit(2)

    from cloverleaf_amd import synth

    if args.copy_probe_mib:
        import torch

        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)  # an elementwise kernel: reads and writes N MiB once
        torch.cuda.synchronize()
        print(f"[bench] copy probe: 3 x {args.copy_probe_mib} MiB elementwise scale (read N MiB, write N MiB)",
              file=sys.stderr, flush=True)
        del src, dst

    t0 = time.time()
    w = synth.make_workload(args.workload, scale=args.scale, graph_scale=args.graph_scale,
                            log=lambda m: print(f"[bench] {m} ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True))
    print(f"[bench] workload {w.name}: {w.n_nodes} nodes, {len(w.indices)} CSR entries, vocab {w.vocab} "
          f"(generated in {time.time() - t0:.1f}s)", file=sys.stderr, flush=True)

    if args.align:
        bench_align(w, args)
        return
    if not args.concurrent:
        args.concurrent = 2 if w.att_dk else 8
    results = {}
    for form in args.forms.split(","):
        best = None
        plain = form.endswith("-plain")
        for sc in args.scatters.split(","):
            if sc == "atomic" and not plain and args.concurrent > 1:
                continue  # the atomic path runs one batch per step
            r = run_form(w, form, args.steps, args.warmup, scatter=sc, concurrent=args.concurrent,
                         graph=not args.no_graph, fp16_adjoints=bool(args.fp16_adjoints))
            ph = " ".join(f"{k}={v * 1e3:.0f}" for k, v in r["phase_ms"].items())
            print(f"[bench] {form}/{sc}: {r['examples_per_s']:.0f} ex/s, {r['ms_per_step']:.3f} ms/step "
                  f"(P {r['concurrent_batches']}, adj {r['adjoint_rows']}, tile {r['seg_tile']}, {r['unique_rows_per_step']:.0f} rows/step, "
                  f"{r['short_rows_per_step']:.0f} short, {r['long_rows_last_step']} long, "
                  f"{r['pieces_last_step']} pieces); "
                  f"phases us: {ph}; fwd {r['forward_gbs']:.0f} GB/s, update {r['update_gbs']:.0f} GB/s, "
                  f"step {r['step_gbs']:.0f} GB/s"
                  + (f", attention {r['attention_tflops']:.1f} TFLOP/s" if w.att_dk else ""),
                  file=sys.stderr, flush=True)
            if best is None or r["examples_per_s"] > best["examples_per_s"]:
                best = r
        if best is not None:
            results[form] = best

    head_form = next((f for f in ("bf16-master-batched", "bf16-batched") if f in results), list(results)[0])
    head = results[head_form]
    cpu = None if args.no_cpu else cpu_baseline(w, args.cpu_budget)
    ph = head["phase_ms"]
    roof = None
    if w.att_dk and ph["forward"] >= ph["short_update"]:
        # the attention forward + backward (scores, softmax, weighted values, their
        # adjoints) on matrix cores and vector ALUs, against the FP32 dense peak
        tf = head["attention_tflops"]
        mm = ("score tiles on BF16 MFMA, dQ / dK on FP32 MFMA, softmax and values on FP32 vector ALUs; "
              "priced at the FP32 peak" if head_form.startswith("bf16") else "FP32 MFMA and vector ALUs")
        roof = {"bound": "mfma", "kernel": "k_att_forward + k_att_loss + k_att_backward (attention blocks: "
                                           f"scores, softmax, values and their adjoints; {mm})",
                "achieved": round(tf, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tf / FP32_PEAK_TFLOPS, 4), "traffic": None,
                "algorithmic_flops_per_launch": round(head["attention_flops_per_step"])}
    if ph["short_update"] >= ph["forward"]:
        kname, kdesc, kbytes, kgbs = ("k_seg_update", "short-row update: reduce runs of adjoint rows + Adam "
                                      "(theta/m/v read and written once per row)",
                                      head["update_bytes_per_launch"], head["update_gbs"])
    else:
        kname, kdesc, kbytes, kgbs = ("k_forward", "gather + mean + margin loss + adjoint",
                                      head["forward_bytes_per_launch"], head["forward_gbs"])
    line = {
        "metric": "EP train examples/sec (fwd+bwd+update), BF16-batched vs FP32 plain, matched AUC",
        "value": round(head["examples_per_s"], 1),
        "unit": "examples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16" if head_form.startswith("bf16") else "fp32",
        "data": "synthetic (seeded Chung-Lu graph + Zipf features, random-init table)",
        "config": {
            "workload": w.name,
            "description": w.description,
            "nodes": w.n_nodes, "csr_entries": int(len(w.indices)), "vocab": w.vocab,
            "dims": w.dims, "batch": w.batch, "negatives": w.negatives, "max_nodes": w.max_nodes,
            **({"attention_heads": w.att_heads, "attention_dk": w.att_dk} if w.att_dk else {}),
            "form": head_form, "concurrent_batches": head["concurrent_batches"], "accumulate": "fp32",
            "forms": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                      for k, v in results.items()},
        },
        "roofline": roof or {
            "bound": "hbm",
            "kernel": f"{kname} ({kdesc})",
            "achieved": round(kgbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(kgbs / HBM_PEAK_GBS, 4),
            "traffic": pmc_traffic(kname, w.name, head_form),
            "algorithmic_bytes_per_launch": round(kbytes),
        },
        "roofline_gather": {
            "bound": "hbm",
            "kernel": "k_forward (gather + mean + loss + adjoint + key emission): the distinct feature rows of "
                      "the launch's examples, their ids and node offsets, over its event-timed duration",
            "achieved": round(head["gather_unique_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(head["gather_unique_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": pmc_traffic("k_forward", w.name, head_form),
            "algorithmic_bytes_per_launch": round(head["gather_unique_bytes_per_launch"]),
            "unique_rows_per_launch": head["gather_unique_rows_per_launch"],
            "occurrences_per_launch": head["gather_occurrences_per_launch"],
        },
        "roofline_step": {
            "bound": "hbm",
            "achieved": round(head["step_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(head["step_gbs"] / HBM_PEAK_GBS, 4),
            "bytes_per_step": round(head["step_bytes"]),
        },
        "cpu_baseline": cpu,
    }
    if "fp32-plain" in results and head_form != "fp32-plain":
        line["config"]["speedup_vs_fp32_plain"] = round(head["examples_per_s"] / results["fp32-plain"]["examples_per_s"], 3)
    if "fp32-batched" in results and head_form != "fp32-batched":
        line["config"]["speedup_vs_fp32_batched"] = round(
            head["examples_per_s"] / results["fp32-batched"]["examples_per_s"], 3)
    print(json.dumps(line))


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Throughput benchmark of the banded seed-extension path on one MI355X.

One step = one pass of the GPU path over the whole batch, with the 2-bit
packed sequences already resident in HBM.  Prints one JSON line.

Workloads (seeded, synthetic; scoring A=1 B=4 O=6 E=1 w=100 zdrop=100
end_bonus=5):
  illumina_150bp_1m  1,048,576 right-extension pairs from 150 bp reads
                     (substitutions 0.2% -> 2% along the read, 0.1% indels),
                     16-bit scores, length-sorted packing: the plain kernel
  uneven_mix_4m      4,194,304 pairs, reads {100,151,250} bp, 5% random
                     targets, 10% w=200
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X public peaks
VALU_INT32_TOPS = 256 * 64 * 2.4e9 / 1e12   # 39.3 T int32 lane-ops/s
OPS_PER_CELL = 16                          # integer ops per DP cell of ksw_extend2's inner loop


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(ps, sw, budget_s=15.0):
    """The CPU oracle (C restatement of ksw_extend2, all host cores) on a
    bounded sample of the same workload."""
    from tests import _oracle as O
    threads = os.cpu_count() or 1
    n = min(ps.n, 16384)
    t0 = time.perf_counter()
    sub = ps.subset(np.arange(n))
    _, cells, _ = O.extend_batch(sw, *sub.flat(), nthreads=threads)
    dt = time.perf_counter() - t0
    # grow the sample to roughly budget_s of CPU work
    target = int(min(ps.n, n * max(1.0, budget_s / max(dt, 1e-3))))
    if target > n:
        sub = ps.subset(np.arange(target))
        t0 = time.perf_counter()
        _, cells, _ = O.extend_batch(sw, *sub.flat(), nthreads=threads)
        dt = time.perf_counter() - t0
        n = target
    return {"value": n / dt, "unit": "pairs/s", "gcups": float(cells.sum()) / dt / 1e9, "cores": threads,
            "kind": "port",
            "sample": "first %d pairs of the workload, C restatement of ksw_extend2 (ksw_affine_gap.rs:952-1169), "
                      "%d host threads, %.1f s" % (n, threads, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="illumina_150bp_1m")
    ap.add_argument("--lanes", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the extension pairs shard independently across GPUs "
              "(no collective on the path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from ferrousalign_amd import synth
    from ferrousalign_amd.banded_swa import DeviceBatch, Context

    t0 = time.perf_counter()
    ps = synth.workload(args.workload)
    sw = synth.default_scoring()
    log("generated %d pairs in %.1f s" % (ps.n, time.perf_counter() - t0))
    ctx = Context(0)
    db = DeviceBatch(sw, *ps.flat(), ctx=ctx)
    db.set_launch(args.lanes)
    flags = 0
    for _ in range(args.warmup):
        db.run(flags)
    log("warmup done")
    t0 = time.perf_counter()
    ms_ev, _, nlaunch = db.time(flags, args.steps)   # K steps, synchronised at both ends
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    out, cells = db.download()
    total_cells = float(cells.astype(np.float64).sum())
    pairs_s = ps.n / (ms_step / 1e3)
    gcups = total_cells / (ms_step / 1e3) / 1e9
    digest = hashlib.sha256(out.tobytes() + cells.astype(np.uint32).tobytes()).hexdigest()
    golden = None
    gpath = os.path.join(ROOT, "tests", "golden", "digests.json")
    if os.path.exists(gpath):
        golden = json.load(open(gpath)).get(args.workload)
    bit_exact = None if golden is None else (golden == digest)

    # dominant kernel launch: HIP events around each launch on the batch's stream
    launches = db.profile_launches(flags)
    order = db.order(flags)
    dom = max(launches, key=lambda x: x["ms"])
    dom_cells = float(cells[order[dom["first"]:dom["first"] + dom["count"]]].astype(np.float64).sum())
    dom_tops = dom_cells * OPS_PER_CELL / (dom["ms"] / 1e3) / 1e12
    roofline = {"bound": "valu", "achieved": round(dom_tops, 3), "peak": round(VALU_INT32_TOPS, 1),
                "unit": "TOPS", "frac": round(dom_tops / VALU_INT32_TOPS, 4), "traffic": None,
                "kernel": "extend_plain_kernel<%d,%d>" % (args.lanes, dom["slots"]),
                "launch_ms": round(dom["ms"], 4), "launch_cells": dom_cells}
    res = {
        "metric": "extension pairs/s (ksw_extend2-exact banded seed extension)",
        "value": round(pairs_s, 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "ms_per_step_events": round(ms_ev, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": args.workload, "pairs": ps.n, "lanes_per_pair": args.lanes,
                   "packing": "length-sorted", "launches_per_step": nlaunch},
        "gcups": round(gcups, 2),
        "bit_exact_vs_golden": bit_exact,
        "digest": digest,
        "roofline": roofline,
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(ps, sw)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

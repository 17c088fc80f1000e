#!/usr/bin/env python3
"""Throughput benchmark of the banded seed-extension path on one MI355X.

One step = one pass of the GPU path over the whole batch, with the 2-bit
packed sequences already resident in HBM.  Prints one JSON line.

Default workload (uneven_mix_4m): 4,194,304 right-extension pairs, read
lengths {100,151,250} bp (weights .3/.5/.2), substitutions rising 0.2%->2%
along the read, 0.1% indels, 5% non-homologous targets, 10% band w=200;
scoring A=1 B=4 O=6 E=1 w=100 zdrop=100 end_bonus=5.  Measured:
  * value: the learned-packed path -- pairs ordered by the work predictor's
    predicted DP cells inside query-length buckets, launch form (bucket
    launches over 4 streams, or persistent waves on a device work queue),
    lanes per pair and queue chunk chosen by the autotuner, replayed as a
    hipGraph;
  * plain: one lane group per pair, length-sorted packing, single launch;
  * length-sorted: the learned path's launch configuration with the
    predictor replaced by length sorting (what the network gains).
The predictor is trained first, on the GPU with the library's kernels, on
cells that the plain kernel records for a separate 16,777,216-pair training
mix drawn the same way (other seed).  Also reported, never as the value:
the same pairs through the host-buffer entry point
(fa_execute_batch_simd_scoring: validation, packing, PCIe both ways, run,
download).
--workload illumina_150bp_1m times the plain kernel on 1,048,576 pairs from
150 bp reads instead; mate_rescue_2m (16-bit) and mate_rescue_xbyte_2m (8-bit,
the mode the reference's mate rescue picks for 151 bp mates), cigar_regen_1m
and chains_150bp_1m time the mate-rescue kernel, CIGAR regeneration and the
chain-extension path.
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

# MI355X public figures: 256 CUs x 64 int32 lane-ops/clk x 2.4 GHz; packed
# 16-bit VALU instructions (v_pk_*) do two ops per lane
VALU_INT32_TOPS = 256 * 64 * 2.4e9 / 1e12
VALU_INT16X2_TOPS = 2 * VALU_INT32_TOPS
OPS_PER_CELL = 16  # integer ops per DP cell of ksw_extend2's inner loop (ksw_affine_gap.rs:1061-1103)
TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "hbm_traffic.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_reference_simd(ps, sw, cells):
    """The reference's own CPU engine restated (oracle/refsimd_baseline.c: its
    SIMD batch path, dispatch.rs:95-268 with kernel.rs / kernel_i16.rs, every
    chunk scored, all host threads, AVX-512BW when present else AVX2) on the
    whole workload.  Its results are not ksw_extend2's (0-based ends, no
    gscore, rows and columns capped at 128), so GCUPS uses the oracle's cell
    count of the same pairs, as the GPU's does."""
    from tests import _oracle as O
    threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    r = O.refsimd_batch(sw, *ps.flat(), isa="auto", nthreads=threads)
    dt = time.perf_counter() - t0
    if r is None:
        return None
    return {"value": round(ps.n / dt, 1), "unit": "pairs/s", "gcups": round(float(cells.sum()) / dt / 1e9, 3),
            "cores": threads, "kind": "port", "isa": r[1], "kernel": "i16" if r[2] else "i8",
            "sample": "all %d pairs through the restated SIMD batch path of the reference (%s, %s kernel, "
                      "%d host threads): %.2f s" % (ps.n, r[1], "i16" if r[2] else "i8", threads, dt)}


def cpu_baseline(ps, sw, budget_s=15.0):
    """The CPU oracle (C restatement of ksw_extend2, all host threads) on a
    bounded sample of the same workload."""
    from tests import _oracle as O
    threads = os.cpu_count() or 1
    n = min(ps.n, 32768)
    sub = ps.subset(np.arange(n))
    t0 = time.perf_counter()
    _, cells, _ = O.extend_batch(sw, *sub.flat(), nthreads=threads)
    dt = time.perf_counter() - t0
    target = int(min(ps.n, n * max(1.0, 0.8 * budget_s / max(dt, 1e-3))))
    if target > n:
        n = target
        sub = ps.subset(np.arange(n))
        t0 = time.perf_counter()
        _, cells, _ = O.extend_batch(sw, *sub.flat(), nthreads=threads)
        dt = time.perf_counter() - t0
    return {"value": round(n / dt, 1), "unit": "pairs/s", "gcups": round(float(cells.sum()) / dt / 1e9, 3),
            "cores": threads, "kind": "port",
            "sample": "first %d pairs of the workload through the C restatement of ksw_extend2 "
                      "(ksw_affine_gap.rs:952-1169) on %d host threads: %.1f s" % (n, threads, dt)}


def timed(db, flags, steps, warmup):
    for _ in range(warmup):
        db.run(flags)
    t0 = time.perf_counter()
    ms_ev, _, nlaunch = db.time(flags, steps)  # K runs enqueued back to back, synchronised at both ends
    wall = time.perf_counter() - t0
    return wall * 1e3 / steps, ms_ev, nlaunch


def wall_steps(step, steps, warmup):
    """Mean wall time (ms) of `step()`, each call synchronous, after `warmup`
    untimed calls: a whole per-batch step, host work included."""
    for _ in range(warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    return (time.perf_counter() - t0) * 1e3 / steps


def traffic_of(kernel, cells):
    """HBM bytes per launch of `kernel` over `cells` DP cells, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (rows 'traffic: <kernel>
    cells=<n> bytes=<b>'), or None when that launch was not profiled."""
    try:
        with open(TRAFFIC_TABLE) as f:
            for line in f:
                p = line.split()
                if len(p) == 4 and p[0] == "traffic:" and p[1] == kernel and p[2] == "cells=%d" % cells:
                    return int(p[3].split("=")[1])
    except OSError:
        pass
    return None


def roofline(db, flags, cells, ctx=None, reps=3):
    """Dominant kernel: every launch of one run is replayed alone on the
    batch's stream and timed with HIP events (reps times, mean); achieved =
    16 ops x the cells that launch computes / its mean duration, against the
    VALU peak of the arithmetic it runs in (packed 16-bit for extend16_*,
    32-bit otherwise), published and measured on this device."""
    runs = [db.profile_launches(flags) for _ in range(reps)]
    ms = np.mean([[x["ms"] for x in r] for r in runs], axis=0)
    launches = runs[0]
    k = int(np.argmax(ms))
    dom = launches[k]
    order = db.order(flags)
    dom_cells = float(cells[order[dom["first"]:dom["first"] + dom["count"]]].astype(np.float64).sum())
    tops = dom_cells * OPS_PER_CELL / (ms[k] / 1e3) / 1e12
    packed = dom["kernel"].startswith(("extend16", "mate16"))
    peak = VALU_INT16X2_TOPS if packed else VALU_INT32_TOPS
    res = {"bound": "valu", "achieved": round(tops, 3), "peak": round(peak, 1), "unit": "TOPS",
           "frac": round(tops / peak, 4), "traffic": traffic_of(dom["kernel"], int(dom_cells)),
           "kernel": dom["kernel"], "peak_dtype": "int16x2" if packed else "int32",
           "launch_ms": round(float(ms[k]), 4), "launch_cells": dom_cells, "ops_per_cell": OPS_PER_CELL}
    if ctx is not None:
        m = ctx.valu_peak(packed)
        res["peak_measured"] = round(m, 2)
        res["frac_of_measured"] = round(tops / m, 4)
    return res


def digest_of(out, cells):
    return hashlib.sha256(out.tobytes() + cells.astype(np.uint32).tobytes()).hexdigest()


def cpu_reference_kswv(ps, threads, m=262144):
    """The reference's own SIMD mate-rescue path restated (oracle/kswv_baseline.c:
    dispatch_kswv_soa with batch_ksw_align_avx2, 32 jobs per chunk, as
    execute_mate_rescue_batch_simd runs it by default on x86) over the first m
    pairs on all host threads.  Its records are not ksw_align2's (approximate
    tb/qb, its own second-best search), so it is a timing baseline only."""
    from tests import _oracle as O
    sub = ps.subset(np.arange(min(m, ps.n)))
    t0 = time.perf_counter()
    out = O.kswv_batch(*sub.flat(), simd=True, nthreads=threads)
    dt = time.perf_counter() - t0
    if out is None:
        return None
    cells = float(sub.n) * 151 * 1500  # every lane sweeps its chunk's full matrix
    return {"value": round(sub.n / dt, 1), "unit": "pairs/s", "gcups": round(cells / dt / 1e9, 3),
            "cores": threads, "kind": "port", "isa": "avx2",
            "sample": "first %d pairs through the restated kswv AVX2 path (kswv_avx2.rs:93-635, "
                      "dispatch.rs:270-395) on %d host threads: %.2f s" % (sub.n, threads, dt)}


def mate_roofline(db, flags, cells, ctx, reps=3):
    """Mate rescue: the forward launches and the separate XSTART reverse
    launches of one run, each replayed alone (reps times, mean), together:
    achieved = 16 ops x all cells of the run / the summed launch time."""
    runs = [db.profile_launches(flags) for _ in range(reps)]
    fwd = float(np.mean([sum(x["ms"] for x in r if x["kernel"].startswith("mate16_kernel")) for r in runs]))
    rev = float(np.mean([sum(x["ms"] for x in r if x["kernel"].startswith("mate16_rev")) for r in runs]))
    ms = fwd + rev
    tops = cells * OPS_PER_CELL / (ms / 1e3) / 1e12
    res = {"bound": "valu", "achieved": round(tops, 3), "peak": round(VALU_INT16X2_TOPS, 1), "unit": "TOPS",
           "frac": round(tops / VALU_INT16X2_TOPS, 4), "traffic": None,
           "kernel": "mate16_kernel + mate16_rev_kernel (every launch of one run)", "peak_dtype": "int16x2",
           "launch_ms": round(ms, 4), "forward_ms": round(fwd, 4), "reverse_ms": round(rev, 4),
           "launches": len(runs[0]), "launch_cells": cells, "ops_per_cell": OPS_PER_CELL}
    m = ctx.valu_peak(True)
    res["peak_measured"] = round(m, 2)
    res["frac_of_measured"] = round(tops / m, 4)
    return res


def mate_bench(args, ctx, golden, byte_mode=False):
    """Mate-rescue local SW (ksw_align2, 16-bit mode; byte_mode: the 8-bit
    mode, xtra | KSW_XBYTE, which mate_rescue.rs:314-321 selects for 151 bp
    mates) on 2,097,152 pairs of a 151 bp query against a 1,500 bp window."""
    from ferrousalign_amd import synth, _native as N
    from ferrousalign_amd.banded_swa import DeviceBatch
    from tests import _oracle as O
    ps = synth.mate_pairs()
    n = ps.n
    xtra = synth.MATE_XTRA | (O.KSW_XBYTE if byte_mode else 0)
    db = DeviceBatch(synth.default_scoring(), *ps.flat(), np.ones(n, np.int32), np.zeros(n, np.int32), ctx=ctx)
    db.set_align2(synth.MATE_MATRIX, 6, 1, 6, 1, xtra)
    flags = N.FA_RUN_ALIGN2
    steps = max(1, min(args.steps, 3))
    ms, ms_ev, nl = timed(db, flags, steps, 1)
    out, cells = db.download_align2()
    total = float(cells.astype(np.float64).sum())
    # the same run with each XSTART pass inside the forward kernel
    ms_fused, _, _ = timed(db, flags | N.FA_ALIGN2_FUSED, steps, 1)
    out_f, cells_f = db.download_align2()
    same_fused = bool((out_f == out).all() and (cells_f == cells).all())
    # the same pairs through the host-buffer call (fa_ksw_align2_batch: checks,
    # packing, upload, run, download), the first call untimed
    from ferrousalign_amd.banded_swa import ksw_align2_batch
    hb = []
    for rep in range(3):
        t0 = time.perf_counter()
        hout, hcells = ksw_align2_batch(*ps.flat(), synth.MATE_MATRIX, 6, 1, 6, 1, xtra, ctx=ctx)
        if rep:
            hb.append((time.perf_counter() - t0) * 1e3)
    host_same = bool(np.array_equal(hout, out) and np.array_equal(hcells, cells))
    from tests.golden.make_golden_mate import digest
    gm = {}
    gpath = os.path.join(ROOT, "tests", "golden", "mate_digests.json")
    if os.path.exists(gpath):
        gm = json.load(open(gpath))
    key = ("mate_u8_%d" if byte_mode else "mate_%d") % n
    exact = (digest(out, cells) == gm[key]) if key in gm else None
    roof = mate_roofline(db, flags, total, ctx)
    mode = "8-bit" if byte_mode else "16-bit"
    res = {"metric": "mate-rescue pairs/s (ksw_align2 %s, bit-exact)" % mode, "value": round(n / (ms / 1e3), 1),
           "unit": "pairs/s", "n_gpus": 1, "steps": steps, "warmup": 1, "ms_per_step": round(ms, 3),
           "ms_per_step_hip_events": round(ms_ev, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u8 in int16x2 halves" if byte_mode else "int16x2 (saturating)",
           "data": "synthetic",
           "config": {"workload": "mate_rescue_xbyte_2m" if byte_mode else "mate_rescue_2m", "pairs": n,
                      "qlen": 151, "tlen": 1500, "xtra": hex(xtra), "launches_per_step": nl},
           "gcups": round(total / (ms / 1e3) / 1e9, 2), "bit_exact_vs_golden": exact,
           "fused_reverse_ms_per_step": round(ms_fused, 3), "identical_to_fused": same_fused,
           "host_buffer_ms": round(float(np.mean(hb)), 2),
           "host_buffer_pairs_per_s": round(n / (float(np.mean(hb)) / 1e3), 1),
           "host_buffer_identical": host_same,
           "roofline": roof}
    if not args.no_cpu:
        threads = os.cpu_count() or 1
        m = 16384
        sub = ps.subset(np.arange(m))
        t0 = time.perf_counter()
        _, oc, _ = O.ksw_align2_batch(*sub.flat(), synth.MATE_MATRIX, 6, 1, 6, 1, xtra, nthreads=threads)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(m / dt, 1), "unit": "pairs/s", "gcups": round(float(oc.sum()) / dt / 1e9, 3),
                               "cores": threads, "kind": "port",
                               "sample": "first %d pairs through the C restatement of ksw_align2 %s "
                                         "(ksw_affine_gap.rs:267-861) on %d host threads: %.1f s"
                                         % (m, mode, threads, dt)}
        res["cpu_baseline_reference_simd"] = cpu_reference_kswv(ps, threads)
    print(json.dumps(res), flush=True)


def cigar_bench(args, ctx):
    """CIGAR regeneration (batched scalar_banded_swa + traceback) on
    1,048,576 region-like pairs: 150 bp reads against the reference span of
    their alignment region."""
    from ferrousalign_amd import synth
    from ferrousalign_amd.banded_swa import (scalar_banded_swa_batch, scalar_banded_swa_batch_dense,
                                             scalar_banded_swa_last_ms)
    from tests import _oracle as O
    flat = synth.region_pairs()
    n = len(flat[2])
    sw = synth.default_scoring()
    h0 = np.zeros(n, np.int32)
    steps = max(1, min(args.steps, 5))
    res_raw = None
    for _ in range(max(1, args.warmup)):
        res_raw = scalar_banded_swa_batch(sw, *flat, h0, ctx=ctx, raw=True)
    dp, tb, wall = [], [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        res_raw = scalar_banded_swa_batch(sw, *flat, h0, ctx=ctx, raw=True)
        wall.append((time.perf_counter() - t0) * 1e3)
        a, b = scalar_banded_swa_last_ms(ctx)
        dp.append(a)
        tb.append(b)
    out, cells, cig, coff, nc = res_raw[:5]
    # the dense entry point: same work, outputs packed back to back
    dense_wall = []
    for _ in range(steps):
        t0 = time.perf_counter()
        dres = scalar_banded_swa_batch_dense(sw, *flat, h0, ctx=ctx)
        dense_wall.append((time.perf_counter() - t0) * 1e3)
    dense_same = bool(np.array_equal(dres[0], out) and np.array_equal(np.diff(dres[3]), nc))
    ms_k = float(np.mean(dp) + np.mean(tb))
    total_cells = float(cells.astype(np.float64).sum())
    # bit-exactness: records, cell counts and CIGAR lengths of every pair against the
    # oracle, full CIGARs and aligned bases of the first 4,096 pairs
    ora, ocells, onc = O.scalar_banded_swa_batch(sw, *flat, h0)
    exact = bool(np.array_equal(ora, out) and np.array_equal(ocells.astype(np.uint32), cells)
                 and np.array_equal(onc, nc))
    m = 4096
    sub = scalar_banded_swa_batch(sw, flat[0], flat[1][:m], flat[2][:m], flat[3], flat[4][:m], flat[5][:m],
                                  flat[6][:m], h0[:m], ctx=ctx)
    for k in range(m):
        q = flat[0][flat[1][k]:flat[1][k] + flat[2][k]]
        t = flat[3][flat[4][k]:flat[4][k] + flat[5][k]]
        rec, c2, r2, q2, _ = O.scalar_banded_swa(q, t, sw, int(flat[6][k]), 0)
        if sub[1][k] != c2 or not np.array_equal(sub[2][k], r2) or not np.array_equal(sub[3][k], q2):
            exact = False
            break
    dp_tops = total_cells * OPS_PER_CELL / (np.mean(dp) / 1e3) / 1e12
    res = {"metric": "CIGAR-regeneration pairs/s (scalar_banded_swa + traceback, bit-exact)",
           "value": round(n / (ms_k / 1e3), 1), "unit": "pairs/s", "n_gpus": 1, "steps": steps,
           "warmup": max(1, args.warmup), "ms_per_step": round(ms_k, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
           "config": {"workload": "cigar_regen_1m", "pairs": n, "qlen": 150, "band": "|tlen-qlen|+10", "h0": 0},
           "gcups": round(total_cells / (ms_k / 1e3) / 1e9, 2),
           "dp_ms": round(float(np.mean(dp)), 4), "traceback_ms": round(float(np.mean(tb)), 4),
           "end_to_end_ms": round(float(np.mean(wall)), 2),
           "end_to_end_pairs_per_s": round(n / (np.mean(wall) / 1e3), 1),
           "end_to_end_dense_ms": round(float(np.mean(dense_wall)), 2),
           "end_to_end_dense_pairs_per_s": round(n / (np.mean(dense_wall) / 1e3), 1),
           "dense_identical": dense_same,
           "bit_exact_vs_oracle": exact,
           "roofline": {"bound": "valu", "achieved": round(dp_tops, 3), "peak": round(VALU_INT32_TOPS, 1),
                        "unit": "TOPS", "frac": round(dp_tops / VALU_INT32_TOPS, 4), "traffic": None,
                        "kernel": "swtb_dp_kernel", "peak_dtype": "int32", "launch_ms": round(float(np.mean(dp)), 4),
                        "launch_cells": total_cells, "ops_per_cell": OPS_PER_CELL,
                        "peak_measured": round(ctx.valu_peak(False), 2)}}
    if not args.no_cpu:
        threads = os.cpu_count() or 1
        t0 = time.perf_counter()
        O.scalar_banded_swa_batch(sw, *flat, h0, nthreads=threads)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(n / dt, 1), "unit": "pairs/s",
                               "gcups": round(total_cells / dt / 1e9, 3), "cores": threads, "kind": "port",
                               "sample": "all %d pairs through the C restatement of scalar_banded_swa "
                                         "(implementation.rs:5-440) on %d host threads: %.1f s" % (n, threads, dt)}
    print(json.dumps(res), flush=True)


def chains_bench(args, ctx):
    """Chain extension (producer, left and right ksw_extend2 batches, region
    merge) on 1,048,576 reads of 150 bp, all on the GPU: targets cut from the
    resident .pac, right h0 chained from the left score, one region per
    chain.  value: reads, chains and seeds resident in HBM
    (fa_chains_extend); also reported: the same through fa_extend_chains
    over host buffers (upload, run, region download)."""
    from ferrousalign_amd import synth, pack_pac
    from ferrousalign_amd.banded_swa import DeviceBatch, DeviceReference
    from ferrousalign_amd.pipeline import (ChainBatch, DeviceChains, MemOpt, collect_extension_jobs,
                                           extend_chains, CHAIN_STAGES)
    from tests import pipeline_oracle as PO
    cb, qseq, qoff, _ = synth.chain_reads()
    ref_codes = synth.reference()
    l_pac = ref_codes.size
    pac = pack_pac(ref_codes)
    ref = DeviceReference(pac, l_pac, ctx=ctx)
    sw = synth.default_scoring()
    opt = MemOpt()
    steps = max(1, args.steps)
    warmup = max(1, args.warmup)
    dc = DeviceChains(qseq, qoff, cb, ctx=ctx)
    for _ in range(warmup):
        nl, nr = dc.extend(ref, sw, opt)
    wall, stages = [], []
    for _ in range(steps):
        st = {}
        t0 = time.perf_counter()
        nl, nr = dc.extend(ref, sw, opt, stage_ms=st)  # returns after the merge has finished
        wall.append((time.perf_counter() - t0) * 1e3)
        stages.append(st)
    ms = float(np.mean(wall))
    stage_ms = {k: round(float(np.mean([x[k] for x in stages])), 3) for k in CHAIN_STAGES[1:7]}
    regions = dc.download()
    dc.free()
    hb_wall, hb_stages = [], []
    for _ in range(max(1, min(steps, 3))):
        st = {}
        t0 = time.perf_counter()
        hregions, _, _ = extend_chains(ref, sw, opt, qseq, qoff, cb, ctx=ctx, stage_ms=st)
        hb_wall.append((time.perf_counter() - t0) * 1e3)
        hb_stages.append(st)
    hb_ms = float(np.mean(hb_wall))
    same_hb = bool(np.array_equal(hregions.view(np.uint8), regions.view(np.uint8)))

    # dominant kernel: the two batches rebuilt from the same jobs and replayed launch by launch
    left, right, m = collect_extension_jobs(opt, l_pac, cb)

    def batch(jobs, reverse):
        ql = jobs["qlen"].astype(np.int64)
        offs = np.concatenate([[0], np.cumsum(ql)[:-1]])
        local = np.arange(int(ql.sum())) - np.repeat(offs, ql)
        base = np.repeat(qoff[jobs["read_idx"]] + jobs["qbeg"], ql)
        idx = base + (np.repeat(ql, ql) - 1 - local if reverse else local)
        return DeviceBatch.from_reference(sw, qseq[idx], offs, jobs["qlen"], ref, jobs["seg_beg"], jobs["seg_len"],
                                          jobs["seg_rev"], jobs["h0"], jobs["w"], ctx=ctx)
    lb = batch(left, True)
    lb.run(0)
    lout, lcells = lb.download()
    rj = right.copy()
    cnt = m.count[:cb.nchains].astype(np.int64)
    idx = np.repeat(m.start[:cb.nchains] - (np.cumsum(cnt) - cnt), cnt) + np.arange(int(cnt.sum()))
    idx = idx[m.right[idx] >= 0]
    li = m.left[idx]
    rj["h0"][m.right[idx]] = np.where(li >= 0, lout[np.maximum(li, 0), 0],
                                      cb.seeds["len"][m.seed[idx]].astype(np.int64) * opt.a)
    rb = batch(rj, False)
    rb.run(0)
    _, rcells = rb.download()
    roofs = [roofline(lb, 0, lcells, ctx), roofline(rb, 0, rcells, ctx)]
    roof = max(roofs, key=lambda r: r["launch_ms"])
    roof["batch"] = "left" if roof is roofs[0] else "right"
    lb.free()
    rb.free()
    total_cells = float(lcells.astype(np.float64).sum() + rcells.astype(np.float64).sum())

    # bit-exactness: the regions of the first reads against the CPU pipeline
    msub = min(cb.nreads, 131072)
    nch = int(cb.read_chain_start[msub]) if msub < cb.nreads else cb.nchains
    nsd = int(cb.chain_seed_start[nch]) if nch < cb.nchains else len(cb.seeds)
    sub = ChainBatch(cb.query_len[:msub], cb.read_chain_start[:msub], cb.read_chain_count[:msub],
                     cb.chain_kept[:nch], cb.chain_frac_rep[:nch], cb.chain_seed_start[:nch],
                     cb.chain_seed_count[:nch], cb.chain_seeds[:nsd], cb.seeds[:nsd])
    threads = os.cpu_count() or 1
    tcpu = {}
    want, _, _ = PO.extend_chains_cpu(sub, pac, l_pac, qseq, qoff[:msub], sw, opt, nthreads=threads, timing=tcpu)
    exact = bool(np.array_equal(want.view(np.uint8), regions[:nch].view(np.uint8)))
    res = {"metric": "chain-extension reads/s (producer + left/right ksw_extend2 batches + region merge, bit-exact)",
           "value": round(cb.nreads / (ms / 1e3), 1), "unit": "reads/s", "n_gpus": 1, "steps": steps,
           "warmup": warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int16x2", "data": "synthetic",
           "config": {"workload": "chains_150bp_1m", "reads": cb.nreads, "chains": cb.nchains,
                      "seeds": int(len(cb.seeds)), "left_jobs": nl, "right_jobs": nr, "read_len": 150,
                      "w": opt.w, "zdrop": opt.zdrop, "inputs": "resident in HBM"},
           "stage_ms": stage_ms,
           "host_buffer_ms": round(hb_ms, 2), "host_buffer_reads_per_s": round(cb.nreads / (hb_ms / 1e3), 1),
           "host_buffer_stage_ms": {k: round(float(np.mean([x[k] for x in hb_stages])), 3) for k in CHAIN_STAGES},
           "host_buffer_identical": same_hb,
           "gcups": round(total_cells / (ms / 1e3) / 1e9, 2),
           "bit_exact_vs_oracle": exact, "bit_exact_sample_reads": msub, "roofline": roof}
    if not args.no_cpu:
        dt = tcpu["producer"] + tcpu["extend"] + tcpu["merge"]
        res["cpu_baseline"] = {"value": round(msub / dt, 1), "unit": "reads/s", "cores": threads, "kind": "port",
                               "sample": "first %d reads: the library's producer and merge with both extension "
                                         "batches through the C restatement of ksw_extend2 on %d host threads "
                                         "(query gathering and target cutting not timed): %.2f s"
                                         % (msub, threads, dt)}
    ref.free()
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="uneven_mix_4m",
                    choices=["uneven_mix_4m", "illumina_150bp_1m", "mate_rescue_2m", "mate_rescue_xbyte_2m",
                             "cigar_regen_1m", "chains_150bp_1m"])
    ap.add_argument("--train-pairs", type=int, default=16777216)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="with --profile: also run a device copy of this many MiB (counter calibration)")
    ap.add_argument("--profile", action="store_true",
                    help="stop after the roofline measurement (isolated launches); for rocprofv3 runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: extension pairs shard independently across GPUs with no "
              "collective on the path, so only --gpus 1 is run here", file=sys.stderr)
        sys.exit(2)

    from ferrousalign_amd import synth, autotune, _native as N
    from ferrousalign_amd.banded_swa import DeviceBatch, Context
    from ferrousalign_amd.predictor import WorkPredictor, mae

    ctx = Context(0)
    sw = synth.default_scoring()
    golden = {}
    gpath = os.path.join(ROOT, "tests", "golden", "digests.json")
    if os.path.exists(gpath):
        golden = json.load(open(gpath))

    if args.workload in ("mate_rescue_2m", "mate_rescue_xbyte_2m"):
        mate_bench(args, ctx, golden, byte_mode=args.workload == "mate_rescue_xbyte_2m")
        return
    if args.workload == "cigar_regen_1m":
        cigar_bench(args, ctx)
        return
    if args.workload == "chains_150bp_1m":
        chains_bench(args, ctx)
        return

    if args.workload == "illumina_150bp_1m":
        ps = synth.workload(args.workload)
        db = DeviceBatch(sw, *ps.flat(), ctx=ctx)
        # one launch for the whole batch in one 8-lane-group kernel: two pairs per
        # group, the slot count of the longest query (wide when every pair is
        # under the wide-key bound)
        db.set_launch(8)
        flags = N.FA_LAUNCH_SINGLE
        ms, ms_ev, nl = timed(db, flags, args.steps, args.warmup)
        out, cells = db.download()
        res = {"metric": "extension pairs/s (ksw_extend2-exact banded seed extension)",
               "value": round(ps.n / (ms / 1e3), 1), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "int16x2", "data": "synthetic",
               "config": {"workload": args.workload, "pairs": ps.n, "lanes_per_pair": 8,
                          "packing": "length-sorted", "launch": "single"},
               "gcups": round(float(cells.astype(np.float64).sum()) / (ms / 1e3) / 1e9, 2),
               "bit_exact_vs_golden": golden.get(args.workload) == digest_of(out, cells),
               "roofline": roofline(db, flags, cells, ctx)}
        if not args.no_cpu:
            res["cpu_baseline"] = cpu_baseline(ps, sw)
            res["cpu_baseline_reference_simd"] = cpu_reference_simd(ps, sw, cells)
        print(json.dumps(res), flush=True)
        return

    # -- 1. train the work predictor on a separate training mix ---------------------
    t0 = time.perf_counter()
    tr = synth.workload("uneven_mix_train_16m", n=args.train_pairs)
    dbt = DeviceBatch(sw, *tr.flat(), ctx=ctx)
    ntrain = tr.n
    del tr
    dbt.set_launch(8)
    dbt.run(N.FA_LAUNCH_GRAPH)  # records the cell labels
    pred = WorkPredictor(ctx, seed=0)
    t1 = time.perf_counter()
    losses = pred.fit(dbt, epochs=args.epochs, minibatch=16384, lr=3e-3, log=log)
    train_s = time.perf_counter() - t1
    dbt.free()
    log("predictor trained on %d pairs in %.2f s (setup %.1f s)" % (ntrain, train_s, t1 - t0))

    # -- 2. evaluation batch -----------------------------------------------------------
    ps = synth.workload(args.workload)
    db = DeviceBatch(sw, *ps.flat(), ctx=ctx)

    if args.profile:
        # isolated launches of the learned-packed configuration only (for rocprofv3)
        pred.predict(db, set_cost=True)
        db.set_launch(0)
        db.set_streams(1)
        flags = N.FA_PACK_COST_ORDER
        db.run(flags)
        _, pcells = db.download()
        print(json.dumps({"roofline": roofline(db, flags, pcells, ctx)}), flush=True)
        if args.copy_mib:
            # calibration for HBM byte counters: one copy kernel that reads
            # and writes copy_mib MiB (launched twice; the second is timed)
            gbps = ctx.hbm_copy(args.copy_mib << 20)
            print(json.dumps({"calibration_copy_bytes": args.copy_mib << 20, "copy_gbps": round(gbps, 1)}),
                  flush=True)
        return

    # plain GPU kernel: one lane group (16 lanes) per pair, length-sorted, single
    # launch; a step plans the packing (device sort by length) and runs it
    db.set_launch(16)
    db.set_streams(1)
    plain_flags = N.FA_PACK_LENGTH | N.FA_LAUNCH_SINGLE
    plain_steps = max(2, args.steps // 2)
    ms_plain_run, _, _ = timed(db, plain_flags, plain_steps, 1)
    ms_plain = wall_steps(lambda: db.run(plain_flags | N.FA_PLAN_FRESH), plain_steps, 1)
    out_plain, cells = db.download()
    log("plain single-launch: %.2f ms/step (%.2f run only)" % (ms_plain, ms_plain_run))

    # learned packing: the predictions stay on the device as the packing cost
    cpred = pred.predict(db, set_cost=True)
    pred_mae = mae(cpred, cells)
    if args.no_tune:
        best = dict(lanes=8, chunk=0, wpc=0, streams=4, launch=N.FA_LAUNCH_GRAPH)
        db.set_launch(8)
        db.set_streams(4)
        flags = N.FA_PACK_COST_ORDER | N.FA_LAUNCH_GRAPH
    else:
        best, flags, _ = autotune.tune(db, ps.qlen, ps.tlen, ps.w, pack_flags=N.FA_PACK_COST_ORDER, log=log,
                                       use_cache=False)
    # run only: the plan of this batch is cached
    ms_run, ms_ev, nlaunch = timed(db, flags, args.steps, args.warmup)
    out_run, cells_run = db.download()
    roof = roofline(db, flags, cells_run, ctx)

    # the headline step, as a new batch costs it: features computed again, the
    # network's forward pass, the packing plan on the device, the run
    def learned_step():
        pred.predict(db, set_cost=True, fresh_features=True, download=False)
        db.run(flags)
    ms = wall_steps(learned_step, args.steps, args.warmup)
    out, cells2 = db.download()
    total_cells = float(cells2.astype(np.float64).sum())

    # ablation: same launch configuration, length-sorted instead of learned
    # (a step plans the packing again), and the packing planned from the true
    # cell counts (what a perfect predictor could give; run only)
    lflags = (flags & ~0xF0) | N.FA_PACK_LENGTH
    len_steps = max(2, args.steps // 2)
    ms_len_run, _, _ = timed(db, lflags, len_steps, 1)
    ms_len = wall_steps(lambda: db.run(lflags | N.FA_PLAN_FRESH), len_steps, 1)
    out_len, cells_len = db.download()
    db.set_cost(cells2.astype(np.float32))
    ms_perfect_run, _, _ = timed(db, flags, len_steps, 1)
    out_perf, cells_perf = db.download()

    # the same pairs through the host-buffer entry point (validation, 2-bit
    # packing, PCIe both ways, packing plan, run, download): the rate a caller
    # of fa_execute_batch_simd_scoring sees; never the headline value
    from ferrousalign_amd.banded_swa import extend_arrays
    flat = ps.flat()
    extend_arrays(sw, *flat, want_cells=True, ctx=ctx)  # first call pays one-time setup
    t0 = time.perf_counter()
    out_h, cells_h = extend_arrays(sw, *flat, want_cells=True, ctx=ctx)
    ms_host = (time.perf_counter() - t0) * 1e3
    host_same = bool(np.array_equal(out_h, out_plain) and np.array_equal(cells_h, cells))
    del flat

    dg = digest_of(out, cells2)
    exact = (dg == golden.get(args.workload)) if golden.get(args.workload) else None
    same = all(np.array_equal(o, out_plain) and np.array_equal(c, cells)
               for o, c in ((out, cells2), (out_run, cells_run), (out_len, cells_len), (out_perf, cells_perf)))
    res = {
        "metric": "extension pairs/s + GCUPS, learned-packed vs plain GPU kernel, bit-exact",
        "value": round(ps.n / (ms / 1e3), 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16x2",
        "data": "synthetic",
        "config": {"workload": args.workload, "pairs": ps.n,
                   "packing": "learned (descending predicted cells per kernel); a step = features + forward "
                              "pass + device packing "
                              "plan + run",
                   "lanes_per_pair": best["lanes"] or "narrowest group per query length",
                   "chunk_pairs": best["chunk"], "streams": best.get("streams", 4),
                   "launch": ("persistent queue" if flags & N.FA_LAUNCH_PERSISTENT else "bucket launches")
                   + " + hipGraph", "launches_per_step": nlaunch},
        "gcups": round(total_cells / (ms / 1e3) / 1e9, 2),
        "plain_pairs_per_s": round(ps.n / (ms_plain / 1e3), 1),
        "plain_ms_per_step": round(ms_plain, 4),
        "length_sorted_ms_per_step": round(ms_len, 4),
        "plain_gcups": round(total_cells / (ms_plain / 1e3) / 1e9, 2),
        "speedup_vs_plain": round(ms_plain / ms, 3),
        "length_sorted_pairs_per_s": round(ps.n / (ms_len / 1e3), 1),
        "learned_gain_vs_length_sorted": round(ms_len / ms, 3),
        "learned_path_ms": round(ms, 4),
        "run_only_ms": {"learned": round(ms_run, 4), "learned_hip_events": round(ms_ev, 4),
                        "length_sorted": round(ms_len_run, 4), "plain": round(ms_plain_run, 4),
                        "perfect_costs": round(ms_perfect_run, 4)},
        "perfect_cost_ms": round(ms_perfect_run, 4),
        "learned_gain_run_only": round(ms_len_run / ms_run, 3),
        "bit_exact_vs_golden": exact,
        "identical_across_packings": same,
        "host_buffer_ms": round(ms_host, 1),
        "host_buffer_pairs_per_s": round(ps.n / (ms_host / 1e3), 1),
        "host_buffer_identical": host_same,
        "digest": dg,
        "predictor": {"train_pairs": ntrain, "epochs": args.epochs, "train_s": round(train_s, 2),
                      "final_train_loss": round(losses[-1], 4), "heldout_mae_cells": round(pred_mae, 1),
                      "mean_cells": round(float(cells.mean()), 1)},
        "roofline": roof,
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(ps, sw)
        res["cpu_baseline_reference_simd"] = cpu_reference_simd(ps, sw, cells)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

"""Launch autotuner for the extension path.

Searches lanes per pair (4, 8, 16, or the narrowest group per query
length), persistent-queue chunk size and launch
form (one launch per query-length bucket, or persistent waves pulling from a
device work counter with the launch sequence replayed as a hipGraph) by
timing the actual batch on the device, and caches the winner per workload
signature in a small JSON file ($FA_AUTOTUNE_CACHE, default
~/.cache/ferrousalign_amd/autotune.json) so later runs skip the search.
Tuning never changes results: every candidate computes the same records.
"""
import hashlib
import json
import os

import numpy as np

from . import _native as N

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "ferrousalign_amd", "autotune.json")


def signature(qlen, tlen, w, pack_flags, n):
    h = hashlib.sha256()
    h.update(np.bincount(np.minimum(qlen, 255), minlength=256).astype(np.int64).tobytes())
    h.update(np.bincount(np.minimum(tlen, 2048) // 16, minlength=129).astype(np.int64).tobytes())
    h.update(np.bincount(np.minimum(w, 1023), minlength=1024).astype(np.int64).tobytes())
    h.update(("%d:%d" % (pack_flags, n)).encode())
    return h.hexdigest()[:24]


def candidates(max_qlen):
    out = []
    for lanes in (0, 16, 8, 4):  # 0: narrowest lane group per query length
        if lanes and (lanes * 32 < max_qlen + 1 or (lanes == 16 and 16 * 16 < max_qlen + 1)):
            continue
        g = 64 // (lanes or 4)
        for streams in (1, 4):
            out.append(dict(lanes=lanes, chunk=0, wpc=0, streams=streams, launch=N.FA_LAUNCH_GRAPH))
        for chunk_tasks in (1, 4):
            out.append(dict(lanes=lanes, chunk=chunk_tasks * g, wpc=0, streams=4,
                            launch=N.FA_LAUNCH_PERSISTENT | N.FA_LAUNCH_GRAPH))
    return out


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, cache):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def tune(db, qlen, tlen, w, pack_flags=N.FA_PACK_LENGTH, reps=2, cache_path=None, log=None, use_cache=True):
    """Pick the fastest launch configuration for DeviceBatch `db`; applies it
    to `db` and returns (config, flags, table of timings)."""
    cache_path = cache_path or os.environ.get("FA_AUTOTUNE_CACHE", DEFAULT_CACHE)
    sig = signature(np.asarray(qlen), np.asarray(tlen), np.asarray(w), pack_flags, len(qlen))
    cache = _load(cache_path) if use_cache else {}
    table = []
    if sig in cache:
        best = cache[sig]
    else:
        best, best_ms = None, float("inf")
        for c in candidates(int(np.max(qlen))):
            db.set_launch(c["lanes"], c["chunk"], c["wpc"])
            db.set_streams(c["streams"])
            flags = pack_flags | c["launch"]
            db.run(flags)  # packing, graph capture, warm caches
            ms, _, _ = db.time(flags, reps)
            table.append(dict(c, ms=ms))
            if log:
                log("autotune lanes=%d chunk=%d streams=%d launch=%s: %.3f ms" % (
                    c["lanes"], c["chunk"], c["streams"],
                    "persistent+graph" if c["launch"] & N.FA_LAUNCH_PERSISTENT else "plain+graph", ms))
            if ms < best_ms:
                best, best_ms = dict(c), ms
        best["ms"] = best_ms
        cache[sig] = best
        if use_cache:
            _store(cache_path, cache)
    db.set_launch(best["lanes"], best["chunk"], best["wpc"])
    db.set_streams(best.get("streams", 4))
    return best, pack_flags | best["launch"], table

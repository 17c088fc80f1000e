"""Writes the golden fixtures of the extension path from the CPU oracle.

  digests.json     SHA-256 over the oracle's full output of each seeded
                   workload: OutScore records (int32 x6, field order of
                   banded_swa/types.rs:77-84) followed by the uint32 cell
                   counts, plus a SHA-256 of the generated inputs
  subset_<name>.npz  the first 256 pairs of each workload with their inputs
                   and the oracle's outputs

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from ferrousalign_amd import synth  # noqa: E402
from tests import _oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
FULL = ["cpu_ref_151bp_4k", "illumina_150bp_1m", "uneven_mix_4m"]


def input_digest(ps):
    h = hashlib.sha256()
    for a in (ps.qlen, ps.tlen, ps.h0, ps.w):
        h.update(np.ascontiguousarray(a, np.int32).tobytes())
    for k in range(ps.n):
        h.update(ps.q[k, :ps.qlen[k]].tobytes())
        h.update(ps.t[k, :ps.tlen[k]].tobytes())
    return h.hexdigest()


def output_digest(out_outscore, cells):
    return hashlib.sha256(np.ascontiguousarray(out_outscore, np.int32).tobytes()
                          + np.ascontiguousarray(cells, np.uint32).tobytes()).hexdigest()


def main():
    sw = synth.default_scoring()
    digests = {}
    for name in FULL:
        ps = synth.workload(name)
        out, cells, rc = O.extend_batch(sw, *ps.flat())
        assert (rc == 0).all()
        oo = O.to_outscore_order(out).astype(np.int32)
        digests[name] = output_digest(oo, cells)
        digests[name + ":inputs"] = input_digest(ps)
        digests[name + ":cells"] = int(cells.sum())
        k = 256
        np.savez_compressed(os.path.join(HERE, "subset_%s.npz" % name),
                            q=ps.q[:k], t=ps.t[:k], qlen=ps.qlen[:k], tlen=ps.tlen[:k], h0=ps.h0[:k], w=ps.w[:k],
                            out=oo[:k], cells=cells[:k].astype(np.uint32))
        print(name, digests[name], digests[name + ":cells"], flush=True)
    with open(os.path.join(HERE, "digests.json"), "w") as f:
        json.dump(digests, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()

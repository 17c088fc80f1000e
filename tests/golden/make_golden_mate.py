"""Golden fixtures of the mate-rescue local SW (ksw_align2) from the CPU
oracle: SHA-256 over the Kswr records (int32 x7: score, te, qe, score2, te2,
tb, qb; ksw_affine_gap.rs:51-61) followed by the uint32 cell counts, for the
first 262,144 pairs and for all 2,097,152 pairs of the seeded workload, in the
16-bit mode (keys mate_N) and the 8-bit mode (xtra | KSW_XBYTE, keys
mate_u8_N), plus the first 128 pairs of the 16-bit mode as a small fixture.

Run from the repository root:  python tests/golden/make_golden_mate.py [--full]
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


def digest(out, cells):
    return hashlib.sha256(np.ascontiguousarray(out, np.int32).tobytes()
                          + np.ascontiguousarray(cells, np.uint32).tobytes()).hexdigest()


def main():
    full = "--full" in sys.argv
    path = os.path.join(HERE, "mate_digests.json")
    d = json.load(open(path)) if os.path.exists(path) else {}
    sizes = [262144] + ([synth.MATE_PAIRS] if full else [])
    ps = synth.mate_pairs(max(sizes))
    flat = ps.subset(np.arange(max(sizes))).flat()
    for tag, xtra in (("mate_u8_", synth.MATE_XTRA | O.KSW_XBYTE), ("mate_", synth.MATE_XTRA)):
        out, cells, rc = O.ksw_align2_batch(*flat, synth.MATE_MATRIX, 6, 1, 6, 1, xtra)
        assert (rc == 0).all()
        for n in sizes:
            d[tag + "%d" % n] = digest(out[:n], cells[:n])
            d[tag + "%d:cells" % n] = int(cells[:n].sum())
            print(tag, n, d[tag + "%d" % n], flush=True)
    k = 128
    np.savez_compressed(os.path.join(HERE, "subset_mate.npz"), q=ps.q[:k], t=ps.t[:k], out=out[:k],
                        cells=cells[:k].astype(np.uint32))
    with open(path, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()

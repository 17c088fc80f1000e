"""Seeded chains and seeds for the producer/consumer tests (test
infrastructure): reads drawn from a random 2-bit reference on both strands,
with substitutions and sometimes one insertion; seeds are the exact runs of at
least `k` bases on each true diagonal, in the bidirectional coordinates of
get_reference_segment (p >= l_pac: reverse complement); chain 0 of a read
holds its true seeds, a decoy chain holds one seed elsewhere, and some chains
are not kept."""
import numpy as np

from ferrousalign_amd import _native as N
from ferrousalign_amd.banded_swa import pack_pac
from ferrousalign_amd.pipeline import ChainBatch


def bidir(ref, l_pac, p):
    return ref[p] if p < l_pac else 3 - ref[2 * l_pac - 1 - p]


def make(seed=0, nreads=200, l_pac=1 << 16, k=19, sub=0.02, exact=False, single_seed=False):
    """single_seed: chain 0 holds one k-base seed inside the first exact run
    instead of all runs (so both extensions run on exact reads)."""
    rng = np.random.default_rng(seed)
    ref = rng.integers(0, 4, l_pac).astype(np.uint8)
    both = np.concatenate([ref, (3 - ref[::-1]).astype(np.uint8)])  # [0, 2 l_pac)
    queries, qlens, seeds = [], [], []
    read_chain_start, read_chain_count = [], []
    chain_kept, chain_frac, chain_seed_start, chain_seed_count, chain_seeds = [], [], [], [], []
    truth = []
    for r in range(nreads):
        R = int(rng.integers(61, 201))
        rev = bool(rng.random() < 0.5)
        p = int(rng.integers(300, l_pac - R - 300))
        P = p if not rev else 2 * l_pac - p - R
        ins = 0 if exact or rng.random() < 0.6 else int(rng.integers(1, 4))
        x = int(rng.integers(30, R - 30))
        if ins:
            q = np.concatenate([both[P:P + x], rng.integers(0, 4, ins).astype(np.uint8),
                                both[P + x:P + R - ins]])
            diag = [(0, x, P), (x + ins, R, P - ins)]  # (query beg, end, ref pos of query pos 0)
        else:
            q = both[P:P + R].copy()
            diag = [(0, R, P)]
        if not exact:
            m = rng.random(R) < sub
            q[m] = (q[m] + 1 + rng.integers(0, 3, int(m.sum()))) % 4
        truth.append((P, R, rev, ins))
        # exact runs >= k on each diagonal
        rs = []
        for qb, qe, base in diag:
            eq = q[qb:qe] == both[base + qb:base + qe]
            run = 0
            for j in range(qe - qb + 1):
                if j < qe - qb and eq[j]:
                    run += 1
                    continue
                if run >= k:
                    s = qb + j - run
                    rs.append((base + s, s, run))
                run = 0
        read_chain_start.append(len(chain_kept))
        nch = 0
        if single_seed and rs:
            rp, qp, ln = rs[0]
            off = int(rng.integers(1, max(2, ln - k)))
            off = min(off, ln - k)
            rs = [(rp + off, qp + off, k)]
        # chain 0: the true seeds
        chain_seed_start.append(len(chain_seeds))
        for rp, qp, ln in rs:
            chain_seeds.append(len(seeds))
            seeds.append((rp, qp, ln))
        chain_seed_count.append(len(rs))
        chain_kept.append(1)
        chain_frac.append(float(rng.random()))
        nch += 1
        if rng.random() < 0.5:  # decoy chain: one random seed elsewhere
            qp = int(rng.integers(0, R - k))
            rp = int(rng.integers(0, 2 * l_pac - R))
            chain_seed_start.append(len(chain_seeds))
            chain_seeds.append(len(seeds))
            seeds.append((rp, qp, k))
            chain_seed_count.append(1)
            chain_kept.append(int(rng.random() < 0.7))
            chain_frac.append(float(rng.random()))
            nch += 1
        read_chain_count.append(nch)
        queries.append(q)
        qlens.append(R)
    sd = np.zeros(len(seeds), N.SEED_DTYPE)
    for i, (rp, qp, ln) in enumerate(seeds):
        sd[i] = (rp, qp, ln)
    cb = ChainBatch(np.array(qlens, np.int32), np.array(read_chain_start, np.int64),
                    np.array(read_chain_count, np.int32), np.array(chain_kept, np.uint8),
                    np.array(chain_frac, np.float32), np.array(chain_seed_start, np.int64),
                    np.array(chain_seed_count, np.int32), np.array(chain_seeds, np.int64), sd)
    qoff = np.zeros(nreads, np.int64)
    qoff[1:] = np.cumsum(qlens)[:-1]
    qseq = np.concatenate(queries).astype(np.uint8)
    return dict(ref=ref, l_pac=l_pac, pac=pack_pac(ref), both=both, cb=cb, qseq=qseq, qoff=qoff, truth=truth)

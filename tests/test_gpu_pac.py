"""GPU parity of targets read from a device-resident .pac reference
(fa_ref_upload / fa_batch_upload_ref): the extracted targets against the CPU
restatement of get_reference_segment and the left-job reversal
(oracle/pac_oracle.c) on random segments of both strands, the extension
results against the same jobs uploaded with host targets, and the full
4,194,304-pair uneven mix rebuilt from reference positions against its golden
digest."""
import json
import os

import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _random_jobs(rng, l_pac, n, max_len=2048):
    beg = np.zeros(n, np.int64)
    ln = rng.integers(0, max_len + 1, n).astype(np.int32)
    ln[:4] = [0, 1, 15, 16]
    rev = rng.integers(0, 2, n).astype(np.uint8)
    fwd = rng.integers(0, 2, n).astype(bool)
    for k in range(n):
        L = int(ln[k])
        if fwd[k]:
            beg[k] = rng.integers(0, l_pac - L + 1)
        else:
            beg[k] = rng.integers(l_pac, 2 * l_pac - L + 1)
    # segments touching the ends and the strand boundary from both sides
    beg[4], ln[4] = l_pac - 37, 37
    beg[5], ln[5] = l_pac, 53
    beg[6], ln[6] = 2 * l_pac - 64, 64
    beg[7], ln[7] = 0, 33
    return beg, ln, rev


def test_targets_match_oracle_segments():
    from ferrousalign_amd import synth, pack_pac
    from ferrousalign_amd.banded_swa import DeviceBatch, DeviceReference
    rng = np.random.default_rng(5)
    l_pac = 200_003  # not a multiple of 4
    seq = rng.integers(0, 4, l_pac).astype(np.uint8)
    pac = pack_pac(seq)
    ref = DeviceReference(pac, l_pac)
    n = 3000
    beg, ln, rev = _random_jobs(rng, l_pac, n)
    q = rng.integers(0, 4, n * 8).astype(np.uint8)
    qoff = np.arange(n, dtype=np.int64) * 8
    qlen = np.full(n, 8, np.int32)
    print("upload", n, flush=True)
    db = DeviceBatch.from_reference(synth.default_scoring(), q, qoff, qlen, ref, beg, ln, rev,
                                    np.full(n, 10, np.int32), np.full(n, 100, np.int32))
    codes, off = db.download_targets()
    for k in range(n):
        want = O.pac_segment(pac, l_pac, beg[k], ln[k], rev[k])
        got = codes[off[k]:off[k] + ln[k]]
        assert np.array_equal(got, want), (k, int(beg[k]), int(ln[k]), int(rev[k]))
    db.free()
    ref.free()


def test_segment_errors():
    from ferrousalign_amd import synth, pack_pac, _native as N
    from ferrousalign_amd.banded_swa import DeviceBatch, DeviceReference
    seq = np.random.default_rng(1).integers(0, 4, 1000).astype(np.uint8)
    ref = DeviceReference(pack_pac(seq), 1000)
    q = np.zeros(8, np.uint8)
    one = lambda v, dt: np.array([v], dt)
    for b, L in ((990, 20), (1995, 10), (-1, 5)):
        with pytest.raises(N.FaError) as e:
            DeviceBatch.from_reference(synth.default_scoring(), q, one(0, np.int64), one(8, np.int32), ref,
                                       one(b, np.int64), one(L, np.int32), one(0, np.uint8), one(10, np.int32),
                                       one(100, np.int32))
        assert e.value.code == N.FA_ERR_SEGMENT
    ref.free()


def test_extension_from_reference_equals_host_targets():
    from ferrousalign_amd import synth, pack_pac
    from ferrousalign_amd.banded_swa import DeviceBatch, DeviceReference
    rng = np.random.default_rng(11)
    ref_codes = synth.reference()
    pac = pack_pac(ref_codes)
    l_pac = ref_codes.size
    ref = DeviceReference(pac, l_pac)
    ps = synth.workload("uneven_mix_4m", n=8192)
    n = ps.n
    # each job's target on a random strand, read forwards or backwards
    rc = rng.integers(0, 2, n).astype(bool)
    rev = rng.integers(0, 2, n).astype(np.uint8)
    beg = np.where(rc, 2 * l_pac - ps.tstart - ps.tlen, ps.tstart).astype(np.int64)
    tq = np.zeros((n, ps.t.shape[1]), np.uint8)
    for k in range(n):
        tq[k, :ps.tlen[k]] = O.pac_segment(pac, l_pac, beg[k], ps.tlen[k], rev[k])
    sw = synth.default_scoring()
    qflat = ps.flat()
    toff = np.arange(n, dtype=np.int64) * tq.shape[1]
    host = DeviceBatch(sw, qflat[0], qflat[1], qflat[2], tq.reshape(-1), toff, ps.tlen, ps.h0, ps.w)
    host.run()
    out_h, cells_h = host.download()
    dev = DeviceBatch.from_reference(sw, qflat[0], qflat[1], qflat[2], ref, beg, ps.tlen, rev, ps.h0, ps.w)
    dev.run()
    out_d, cells_d = dev.download()
    assert np.array_equal(out_h, out_d) and np.array_equal(cells_h, cells_d)
    # and the host-target batch itself equals the oracle
    ora, oc, orc = O.extend_batch(sw, qflat[0], qflat[1], qflat[2], tq.reshape(-1), toff, ps.tlen, ps.h0, ps.w)
    assert (orc == 0).all()
    assert np.array_equal(O.to_outscore_order(ora), out_d) and np.array_equal(oc.astype(np.uint32), cells_d)
    host.free()
    dev.free()
    ref.free()


def test_full_uneven_mix_from_reference_matches_golden():
    from ferrousalign_amd import synth, pack_pac
    from ferrousalign_amd.banded_swa import DeviceBatch, DeviceReference
    from tests.golden.make_golden import output_digest
    golden = json.load(open(os.path.join(HERE, "golden", "digests.json")))
    ref_codes = synth.reference()
    ref = DeviceReference(pack_pac(ref_codes), ref_codes.size)
    ps = synth.workload("uneven_mix_4m")
    qflat = ps.flat()
    print("upload 4M from reference", flush=True)
    db = DeviceBatch.from_reference(synth.default_scoring(), qflat[0], qflat[1], qflat[2], ref, ps.tstart,
                                    ps.tlen, np.zeros(ps.n, np.uint8), ps.h0, ps.w)
    db.run()
    out, cells = db.download()
    assert output_digest(out, cells) == golden["uneven_mix_4m"]
    db.free()
    ref.free()

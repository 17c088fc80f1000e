"""Second restatement of ksw_align2 in its 16-bit and 8-bit modes (test
only), written as whole-vector numpy operations over the lanes of one SSE
register, the way the reference's SimdEngine128 executes it
(ksw_affine_gap.rs:78-151, :267-494, :503-735, :773-861): 8 signed 16-bit
lanes, or 16 unsigned byte lanes with KSW_XBYTE.  Used to pin the C oracle on
small inputs."""
import numpy as np

P = 8
XSTOP, XSUBO, XSTART, XBYTE = 0x20000, 0x40000, 0x80000, 0x10000


def _sat(x):
    return np.clip(x, -32768, 32767).astype(np.int32)


def _profile(query, m, mat, p=P):
    qlen = len(query)
    slen = -(-qlen // p)
    k = np.arange(slen)[:, None] + np.arange(p)[None, :] * slen          # [slen, p] query positions
    qp = np.zeros((m, slen, p), np.int32)
    valid = k < qlen
    for a in range(m):
        row = np.asarray(mat[a * m:(a + 1) * m], np.int32)
        qp[a][valid] = row[np.asarray(query)[k[valid]]]
    return qp, slen, k


def _u8(query, target, m, mat, o_del, e_del, o_ins, e_ins, xtra):
    """ksw_u8_impl: bytes as ints clipped to [0, 255] after every saturating step."""
    p = 16
    qp, slen, kpos = _profile(query, m, mat, p)
    mn = min(127, int(min(mat)))
    qmax = max(0, int(max(mat)))
    shift = (256 - mn) & 0xFF
    qp = (qp + shift) & 0xFF                      # the biased profile byte, shift past qlen
    u = lambda x: np.clip(x, 0, 255)
    minsc = xtra & 0xffff if xtra & XSUBO else 0x10000
    endsc = xtra & 0xffff if xtra & XSTOP else 0x10000
    oed, ed, oei, ei = (o_del + e_del) & 0xFF, e_del & 0xFF, (o_ins + e_ins) & 0xFF, e_ins & 0xFF
    H0 = np.zeros((slen, p), np.int32)
    E = np.zeros((slen, p), np.int32)
    Hmax = np.zeros((slen, p), np.int32)
    b = []
    gmax, te, rows = 0, -1, 0
    for i in range(len(target)):
        rows += 1
        f = np.zeros(p, np.int32)
        mx = np.zeros(p, np.int32)
        h = np.concatenate([[0], H0[slen - 1, :p - 1]])
        s = qp[target[i]]
        H1 = np.zeros((slen, p), np.int32)
        for j in range(slen):
            h = u(u(h + s[j]) - shift)
            e = E[j].copy()
            h = np.maximum(np.maximum(h, e), f)
            mx = np.maximum(mx, h)
            H1[j] = h
            E[j] = np.maximum(u(e - ed), u(h - oed))
            f = np.maximum(u(f - ei), u(h - oei))
            h = H0[j].copy()
        for _ in range(p):
            f = np.concatenate([[0], f[:p - 1]])
            changed = False
            for j in range(slen):
                hn = np.maximum(H1[j], f)
                changed |= bool((hn != H1[j]).any())
                H1[j] = hn
                f = u(f - ei)
            if not changed:
                break
        imax = int(mx.max())
        if imax >= minsc:
            if not b or b[-1][1] + 1 != i:
                b.append([imax, i])
            elif b[-1][0] < imax:
                b[-1] = [imax, i]
        if imax > gmax:
            gmax, te = imax, i
            Hmax = H1.copy()
            if gmax + shift >= 255 or gmax >= endsc:
                break
        H0 = H1
    score = gmax if gmax + shift < 255 else 255
    r = dict(score=score, te=te, qe=-1, score2=-1, te2=-1, tb=-1, qb=-1)
    if score != 255:
        valid = kpos < len(query)
        vals = np.where(valid, Hmax, -(1 << 20))
        best = vals.max()
        r["qe"] = int(kpos[(vals == best) & valid].min())
        if b:
            iv = (score + qmax - 1) // qmax
            lo, hi = te - iv, te + iv
            for sc, e in b:
                if (e < lo or e > hi) and sc > r["score2"]:
                    r["score2"], r["te2"] = sc, e
    return r, rows


def _i16(query, target, m, mat, o_del, e_del, o_ins, e_ins, xtra):
    qp, slen, kpos = _profile(query, m, mat)
    qmax = max(0, int(max(mat)))
    minsc = xtra & 0xffff if xtra & XSUBO else 0x10000
    endsc = xtra & 0xffff if xtra & XSTOP else 0x10000
    H0 = np.zeros((slen, P), np.int32)
    E = np.zeros((slen, P), np.int32)
    Hmax = np.zeros((slen, P), np.int32)
    b = []
    gmax, te, rows = 0, -1, 0
    for i in range(len(target)):
        rows += 1
        f = np.zeros(P, np.int32)
        mx = np.zeros(P, np.int32)
        h = np.concatenate([[0], H0[slen - 1, :P - 1]])
        s = qp[target[i]]
        H1 = np.zeros((slen, P), np.int32)
        for j in range(slen):
            h = _sat(h + s[j])
            e = E[j].copy()
            h = np.maximum(np.maximum(h, e), f)
            mx = np.maximum(mx, h)
            H1[j] = h
            E[j] = np.maximum(_sat(e - e_del), _sat(h - (o_del + e_del)))
            f = np.maximum(_sat(f - e_ins), _sat(h - (o_ins + e_ins)))
            h = H0[j].copy()
        for _ in range(P):
            f = np.concatenate([[0], f[:P - 1]])
            changed = False
            for j in range(slen):
                hn = np.maximum(H1[j], f)
                changed |= bool((hn != H1[j]).any())
                H1[j] = hn
                f = _sat(f - e_ins)
            if not changed:
                break
        imax = int(mx.max())
        if imax >= minsc:
            if not b or b[-1][1] + 1 != i:
                b.append([imax, i])
            elif b[-1][0] < imax:
                b[-1] = [imax, i]
        if imax > gmax:
            gmax, te = imax, i
            Hmax = H1.copy()
            if gmax >= 0x7FFF or gmax >= endsc:
                break
        H0 = H1
    r = dict(score=gmax, te=te, qe=-1, score2=-1, te2=-1, tb=-1, qb=-1)
    if gmax != 0x7FFF:
        valid = kpos < len(query)
        vals = np.where(valid, Hmax, -(1 << 20))
        best = vals.max()
        r["qe"] = int(kpos[(vals == best) & valid].min())
        if b:
            iv = (gmax + qmax - 1) // qmax
            lo, hi = te - iv, te + iv
            for sc, e in b:
                if (e < lo or e > hi) and sc > r["score2"]:
                    r["score2"], r["te2"] = sc, e
    return r, rows


def ksw_align2_i16(query, target, mat, o_del, e_del, o_ins, e_ins, xtra, m=5):
    """Returns ((score, te, qe, score2, te2, tb, qb), cells); the mode is the
    one xtra selects (8-bit with XBYTE)."""
    query = list(query)
    target = list(target)
    run = _u8 if xtra & XBYTE else _i16
    r, rows = run(query, target, m, mat, o_del, e_del, o_ins, e_ins, xtra)
    cells = rows * len(query)
    find = (xtra & XSTART) and not ((xtra & XSUBO) and r["score"] < (xtra & 0xffff))
    if find and r["qe"] >= 0 and r["te"] >= 0:
        ql, tl = r["qe"] + 1, r["te"] + 1
        qr = query[:ql][::-1]
        tr = target[:tl][::-1] + target[tl:]
        rr, rows2 = run(qr, tr, m, mat, o_del, e_del, o_ins, e_ins, XSTOP | r["score"])
        cells += rows2 * ql
        if rr["score"] >= r["score"]:
            r["tb"] = r["te"] - rr["te"]
            r["qb"] = r["qe"] - rr["qe"]
    return tuple(r[k] for k in ("score", "te", "qe", "score2", "te2", "tb", "qb")), cells


ksw_align2 = ksw_align2_i16

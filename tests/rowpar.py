"""Independent restatement of ksw_extend2 in row-parallel form (test only).

The C oracle follows the reference's sequential column loop
(src/core/alignment/ksw_affine_gap.rs:1056-1104).  This module computes the
same dynamic program the way the GPU kernel does, from whole-row vector
identities instead of the running `f`, `h1`, `m` scalars:

* F(j) = max_{beg<=k<j} (max(M_k - oe_ins, 0) + k*e_ins) - (j-1)*e_ins,
  F(beg) = 0;
* the row maximum and its first column from one max over keys;
* eh[j].h <- H(i, j-1) as a shift of the H row;
* the next band from the first/last non-zero column.

Agreement of the two on exhaustive tiny inputs pins the C transcription and
the parallel reformulation against each other.  Pure Python: small inputs.
"""


def capped_w(qlen, mat, o_del, e_del, o_ins, e_ins, w, end_bonus):
    mx = max(0, max(int(v) for v in mat))

    def as_i32(x):
        if x != x:
            return 0
        if x >= 2147483647.0:
            return 2147483647
        if x <= -2147483648.0:
            return -2147483648
        return int(x)  # truncation toward zero

    def div(a, b):
        if b == 0:
            return float("inf") if a > 0 else (float("-inf") if a < 0 else float("nan"))
        return a / b

    mi = max(as_i32(div(float(qlen * mx + end_bonus - o_ins), float(e_ins)) + 1.0), 1)
    w = min(w, mi)
    md = max(as_i32(div(float(qlen * mx + end_bonus - o_del), float(e_del)) + 1.0), 1)
    return min(w, md)


def extend(query, target, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, h0, m=5, with_rows=False):
    """Returns ((score, qle, tle, gtle, gscore, max_off), cells), and the rows
    the row loop ran (the break row included) when with_rows."""
    qlen, tlen = len(query), len(target)
    assert h0 > 0 and qlen > 0
    oe_del, oe_ins = o_del + e_del, o_ins + e_ins
    # eh state, written as whole vectors
    v1 = h0 - oe_ins if h0 > oe_ins else 0
    Hs = [h0] + [max(v1 - (c - 1) * e_ins, 0) for c in range(1, qlen + 1)]
    Es = [0] * (qlen + 1)
    w = capped_w(qlen, mat, o_del, e_del, o_ins, e_ins, w, end_bonus)
    best, max_i, max_j, max_ie, gscore, max_off = h0, -1, -1, -1, -1, 0
    beg, end = 0, qlen
    cells = 0
    rows = 0
    NEG = -(1 << 40)
    for i in range(tlen):
        rows = i + 1
        rb = max(beg, i - w)
        re = min(end, i + w + 1, qlen)
        cells += max(re - rb, 0)
        h1i = max(h0 - (o_del + e_del * (i + 1)), 0) if rb == 0 else 0
        cols = range(rb, re)
        M = {c: (Hs[c] + mat[target[i] * m + query[c]]) if Hs[c] != 0 else 0 for c in cols}
        G = {c: max(M[c] - oe_ins, 0) + c * e_ins for c in cols}
        H = {}
        run = NEG
        for c in cols:  # exclusive prefix maximum of G
            F = run - (c - 1) * e_ins if c > rb else 0
            H[c] = max(M[c], Es[c], F)
            run = max(run, G[c])
        newE = {c: max(Es[c] - e_del, M[c] - oe_del, 0) for c in cols}
        # row max with the first column on ties: max of (H, -c)
        if cols:
            m_row, negc = max((H[c], -c) for c in cols)
            mj = -negc if m_row > 0 else -1
        else:
            m_row, mj = 0, -1
        # shifted write-back on [rb, re] (eh[rb] takes the first-column value)
        for c in cols:
            Es[c] = newE[c]
        written_h = {}
        for c in cols:
            written_h[c] = h1i if c == rb else H[c - 1]
        written_h[re] = (H[re - 1] if re - 1 >= rb else h1i)
        for c, v in written_h.items():
            Hs[c] = v
        Es[re] = 0
        if re == qlen and gscore < Hs[re]:
            gscore, max_ie = Hs[re], i
        if m_row == 0:
            break
        if m_row > best:
            best, max_i, max_j = m_row, i, mj
            max_off = max(max_off, abs(mj - i))
        elif zdrop > 0:
            di, dj = i - max_i, mj - max_j
            if di > dj:
                if best - m_row - (di - dj) * e_del > zdrop:
                    break
            elif best - m_row - (dj - di) * e_ins > zdrop:
                break
        nz = [c for c in range(rb, re + 1) if Hs[c] != 0 or Es[c] != 0]
        first = [c for c in nz if c < re]
        nb = first[0] if first else re
        ne = max(nb, nz[-1]) if nz else nb
        beg, end = nb, min(ne + 2, qlen)
    res = (best, max_j + 1, max_i + 1, max_ie + 1, gscore, max_off)
    return (res, cells, rows) if with_rows else (res, cells)

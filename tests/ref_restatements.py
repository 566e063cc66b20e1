"""Independent restatements used to pin the C oracle (tests only).

- ksw_extend2_bwa: bwa's published ksw_extend2 (ksw.c, eh_t form), written
  from the published C algorithm rather than from the oracle.
- gotoh_global: textbook full-matrix affine global alignment score (Gotoh),
  separate insertion/deletion penalties; equals GlobalAligner's score whenever
  the band cannot bind.
- rescore_cigar: score of a CIGAR path under affine gaps.
- local_affine_brute: textbook full-matrix local SW score with gap cost
  o + e*(len+1) per run, the gap model of LocalSWAligner's striped pass
  (open from H: satsub(H, o+e), then extend: satsub(., e)).
"""


def ksw_extend2_bwa(query, target, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, h0):
    qlen, tlen, m = len(query), len(target), 5
    if qlen == 0 or tlen == 0:
        return (0, 0, 0, 0, 0, 0)
    oe_del, oe_ins = o_del + e_del, o_ins + e_ins
    qp = [[mat[k * m + query[j]] for j in range(qlen)] for k in range(m)]
    eh_h = [0] * (qlen + 1)
    eh_e = [0] * (qlen + 1)
    eh_h[0] = h0
    eh_h[1] = h0 - oe_ins if h0 > oe_ins else 0
    j = 2
    while j <= qlen and eh_h[j - 1] > e_ins:
        eh_h[j] = eh_h[j - 1] - e_ins
        j += 1
    mx = max([0] + list(mat))
    max_ins = int((qlen * mx + end_bonus - o_ins) / e_ins + 1.0)
    max_ins = max(max_ins, 1)
    w = min(w, max_ins)
    max_del = int((qlen * mx + end_bonus - o_del) / e_del + 1.0)
    max_del = max(max_del, 1)
    w = min(w, max_del)
    best, max_i, max_j, max_ie, gscore, max_off = h0, -1, -1, -1, -1, 0
    beg, end = 0, qlen
    for i in range(tlen):
        f, mrow, mj = 0, 0, -1
        q = qp[target[i]]
        if beg < i - w:
            beg = i - w
        if end > i + w + 1:
            end = i + w + 1
        if end > qlen:
            end = qlen
        h1 = max(h0 - (o_del + e_del * (i + 1)), 0) if beg == 0 else 0
        j = beg
        while j < end:
            M, e = eh_h[j], eh_e[j]
            eh_h[j] = h1
            M = M + q[j] if M else 0
            h = M if M > e else e
            h = h if h > f else f
            h1 = h
            mj = mj if mrow > h else j
            mrow = mrow if mrow > h else h
            t = max(M - oe_del, 0)
            e = max(e - e_del, t)
            eh_e[j] = e
            t = max(M - oe_ins, 0)
            f = max(f - e_ins, t)
            j += 1
        eh_h[end], eh_e[end] = h1, 0
        if j == qlen:
            max_ie = max_ie if gscore > h1 else i
            gscore = gscore if gscore > h1 else h1
        if mrow == 0:
            break
        if mrow > best:
            best, max_i, max_j = mrow, i, mj
            max_off = max(max_off, abs(mj - i))
        elif zdrop > 0:
            if i - max_i > mj - max_j:
                if best - mrow - ((i - max_i) - (mj - max_j)) * e_del > zdrop:
                    break
            else:
                if best - mrow - ((mj - max_j) - (i - max_i)) * e_ins > zdrop:
                    break
        j = beg
        while j < end and eh_h[j] == 0 and eh_e[j] == 0:
            j += 1
        beg = j
        j = end
        while j >= beg and eh_h[j] == 0 and eh_e[j] == 0:
            j -= 1
        end = j + 2 if j + 2 < qlen else qlen
    return (best, max_j + 1, max_i + 1, max_ie + 1, gscore, max_off)


NEG = -(1 << 40)


def gotoh_global(query, target, mat, o_del, e_del, o_ins, e_ins):
    """Optimal global score; rows = target (deletions consume target)."""
    n, m = len(target), len(query)
    H = [[NEG] * (m + 1) for _ in range(n + 1)]
    E = [[NEG] * (m + 1) for _ in range(n + 1)]   # vertical: deletion (consumes target)
    F = [[NEG] * (m + 1) for _ in range(n + 1)]   # horizontal: insertion (consumes query)
    H[0][0] = 0
    for j in range(1, m + 1):
        F[0][j] = H[0][j] = -(o_ins + e_ins * j)
    for i in range(1, n + 1):
        E[i][0] = H[i][0] = -(o_del + e_del * i)
        for j in range(1, m + 1):
            E[i][j] = max(H[i - 1][j] - o_del - e_del, E[i - 1][j] - e_del)
            F[i][j] = max(H[i][j - 1] - o_ins - e_ins, F[i][j - 1] - e_ins)
            H[i][j] = max(H[i - 1][j - 1] + mat[target[i - 1] * 5 + query[j - 1]], E[i][j], F[i][j])
    return H[n][m]


def rescore_cigar(cigar, query, target, mat, o_del, e_del, o_ins, e_ins):
    s, qi, ti = 0, 0, 0
    for c in cigar:
        op, ln = c & 0xF, c >> 4
        if op == 0:
            for _ in range(ln):
                s += mat[target[ti] * 5 + query[qi]]
                qi += 1
                ti += 1
        elif op == 1:
            s -= o_ins + e_ins * ln
            qi += ln
        elif op == 2:
            s -= o_del + e_del * ln
            ti += ln
    return s, qi, ti


def local_affine_brute(query, target, mat, o, e):
    """Best local score with gap-run cost o + e*(len+1); H clamps at 0."""
    n, m = len(target), len(query)
    best = 0
    Hp = [0] * (m + 1)
    Ep = [0] * (m + 1)
    for i in range(1, n + 1):
        Hc = [0] * (m + 1)
        Ec = [0] * (m + 1)
        f = 0
        for j in range(1, m + 1):
            Ec[j] = max(Hp[j] - o - e - e, Ep[j] - e, 0)
            f = max(Hc[j - 1] - o - e - e, f - e, 0)
            h = max(Hp[j - 1] + mat[target[i - 1] * 5 + query[j - 1]], Ec[j], f, 0)
            Hc[j] = h
            best = max(best, h)
        Hp, Ep = Hc, Ec
    return best

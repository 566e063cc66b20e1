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
- local_sw_swift: LocalSWAligner.align (LocalSWAligner.swift:42-492) written
  from the Swift source with explicit 16-lane stripes as Python lists, to
  check the C oracle's transcription (a second, separately written one).
- global_tb_full: banded global alignment as three full (tlen+1) x (qlen+1)
  matrices with the tie policy stated once, declaratively, and a traceback
  that re-derives each step from the matrices (no bit flags, no rolling
  arrays); checks the C oracle's GlobalAligner fill and traceback.
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


# ---------------------------------------------------------------------------
# LocalSWAligner (LocalSWAligner.swift), transcribed lane by lane
# ---------------------------------------------------------------------------
def _satsub(a, b):
    return a - b if a > b else 0


def _striped_pass(query, target, mat, gap_oe, gap_e, bias, min_sub):
    """simdSWPass (:133-306).  None on overflow, else (score, qEnd, tEnd, score2)."""
    L = 16
    qlen, tlen = len(query), len(target)
    sn = (qlen + L - 1) // L
    prof = [[[0] * L for _ in range(sn)] for _ in range(5)]
    for k in range(5):
        for s in range(sn):
            for lane in range(L):
                j = s + lane * sn
                if j < qlen:
                    prof[k][s][lane] = min(255, max(0, mat[k * 5 + query[j]] + bias))
    H = [[0] * L for _ in range(sn)]
    E = [[0] * L for _ in range(sn)]
    goe, ge = gap_oe & 0xFF, gap_e & 0xFF
    shr = lambda v: [0] + v[:L - 1]                     # shiftRight: lane l -> l + 1
    best, bi, bj = 0, -1, -1
    peaks = []                                          # [score, column]
    track = min_sub > 0
    max_base = 0
    if track:
        max_base = max(min(255, max(0, m + bias)) for m in mat[:25])
        max_base = max_base - bias if max_base > bias else 0
    for i in range(tlen):
        tb = min(target[i], 4)
        vH = shr(H[sn - 1])
        f = [0] * L
        for s in range(sn):
            hp = H[s]
            hn = [0] * L
            for lane in range(L):
                x = (vH[lane] + prof[tb][s][lane]) & 0xFF          # wrapping add
                x = _satsub(x, bias)                               # pmax(x, bias) - bias
                ev = _satsub(max(_satsub(hp[lane], goe), E[s][lane]), ge)
                E[s][lane] = ev
                x = max(x, ev, f[lane])
                hn[lane] = x
                f[lane] = _satsub(max(_satsub(x, goe), f[lane]), ge)
            vH = hp
            H[s] = hn
        cf = shr(f)                                     # single lazy-F pass (:234-242)
        for s in range(sn):
            nh = [max(H[s][lane], cf[lane]) for lane in range(L)]
            if nh == H[s]:
                break
            H[s] = nh
            cf = [_satsub(max(_satsub(H[s][lane], goe), cf[lane]), ge) for lane in range(L)]
        col_max = 0
        for s in range(sn):                             # stripes, then lanes (:245-263)
            lm = max(H[s])
            if lm > 250:
                return None
            col_max = max(col_max, lm)
            if lm > best:
                best, bi = lm, i
                for lane in range(L):
                    if H[s][lane] == lm and s + lane * sn < qlen:
                        bj = s + lane * sn
                        break
        if track and col_max >= min_sub:                # peaks (:266-286)
            if not peaks or peaks[-1][1] + 1 != i:
                peaks.append([col_max, i])
            elif peaks[-1][0] < col_max:
                peaks[-1] = [col_max, i]
    score2 = 0
    if track and peaks and best > 0 and max_base > 0:   # :289-303
        r = (best + max_base - 1) // max_base
        lo, hi = bi - r, bi + r
        for sc, col in peaks:
            if (col < lo or col > hi) and sc > score2:
                score2 = sc
    return best, bj, bi, score2


def _scalar_pass(query, target, mat, gap_oe, gap_e, min_sub):
    """scalarSWPass (:383-492): query outer, strict '>' for the best cell."""
    qlen, tlen = len(query), len(target)
    h = [0] * (tlen + 1)
    e = [0] * (tlen + 1)
    best, bq, bt = 0, -1, -1
    peaks = []
    track = min_sub > 0
    max_base = max([0] + list(mat[:25])) if track else 0
    for i in range(qlen):
        qb = min(query[i], 4)
        f = hprev = row_max = 0
        for j in range(tlen):
            tb = min(target[j], 4)
            hc = hprev + mat[qb * 5 + tb]
            f = max(h[j + 1] - gap_oe, f - gap_e)
            ev = max(h[j] - gap_oe, e[j + 1] - gap_e)
            e[j + 1] = ev
            hc = max(hc, f, ev, 0)
            hprev = h[j + 1]
            h[j + 1] = hc
            row_max = max(row_max, hc)
            if hc > best:
                best, bq, bt = hc, i, j
        if track and row_max >= min_sub:
            if not peaks or peaks[-1][1] + 1 != i:
                peaks.append([row_max, i])
            elif peaks[-1][0] < row_max:
                peaks[-1] = [row_max, i]
    score2 = 0
    if track and peaks and best > 0 and max_base > 0:
        r = (best + max_base - 1) // max_base
        lo, hi = bq - r, bq + r
        for sc, col in peaks:
            if (col < lo or col > hi) and sc > score2:
                score2 = sc
    return best, bq, bt, score2


def local_sw_swift(query, target, mat, o_ins, e_ins, a, b, min_sub=0):
    """LocalSWAligner.align (:42-124): (score, qb, qe, tb, te, score2) or None."""
    query, target = list(query), list(target)
    if not query or not target:
        return None
    gap_oe, gap_e = o_ins + e_ins, e_ins
    if len(query) * a <= 250:
        fwd = _striped_pass(query, target, mat, gap_oe, gap_e, b, min_sub)
        if fwd is not None and fwd[0] > 0:
            fs, fq, ft, f2 = fwd
            rev = _striped_pass(query[:fq + 1][::-1], target[:ft + 1][::-1], mat, gap_oe, gap_e, b, 0)
            if rev is not None:
                return (fs, fq - rev[1], fq, ft - rev[2], ft, f2)
    fs, fq, ft, f2 = _scalar_pass(query, target, mat, gap_oe, gap_e, min_sub)
    if fs <= 0:
        return None
    rev = _scalar_pass(query[:fq + 1][::-1], target[:ft + 1][::-1], mat, gap_oe, gap_e, 0)
    return (fs, fq - rev[1], fq, ft - rev[2], ft, f2)


# ---------------------------------------------------------------------------
# GlobalAligner (GlobalAligner.swift:19-232) as full matrices
# ---------------------------------------------------------------------------
def global_tb_full(query, target, mat, o_del, e_del, o_ins, e_ins, w):
    """Banded affine global alignment, rows = target, columns = query.

    Tie policy (what GlobalAligner's strict comparisons amount to): a gap
    state extends rather than opens when both give the same value; H takes
    the diagonal over E and E over F when values tie.  Band: row i holds
    columns max(0, i - w') .. min(q, i + w' + 1) - 1 with w' = max(w, |q - t|);
    cells outside are -inf, the left boundary column only exists while the
    band starts at column 0.  Returns (score, packed CIGAR ops)."""
    INF = float("inf")
    q, t = len(query), len(target)
    if q == 0 or t == 0:
        return 0, []
    w = max(w, abs(q - t))
    # matrices indexed [i + 1][j + 1]; row 0 / column 0 are the boundaries
    H = [[-INF] * (q + 1) for _ in range(t + 1)]
    E = [[-INF] * (q + 1) for _ in range(t + 1)]
    F = [[-INF] * (q + 1) for _ in range(t + 1)]
    H[0][0] = 0
    for j in range(1, min(w, q) + 1):
        H[0][j] = -(o_ins + e_ins * j)
    inband = lambda i, j: max(0, i - w) <= j < min(q, i + w + 1)
    for i in range(t):
        if i - w <= 0:
            H[i + 1][0] = -(o_del + e_del * (i + 1))
        for j in range(q):
            if not inband(i, j):
                continue
            E[i + 1][j + 1] = max(H[i][j + 1] - o_del - e_del, E[i][j + 1] - e_del)
            F[i + 1][j + 1] = max(H[i + 1][j] - o_ins - e_ins, F[i + 1][j] - e_ins)
            M = H[i][j] + mat[target[i] * 5 + query[j]]
            H[i + 1][j + 1] = max(M, E[i + 1][j + 1], F[i + 1][j + 1])
    score = H[t][q]

    def came_from(i, j):
        M = H[i][j] + mat[target[i] * 5 + query[j]]
        e, f = E[i + 1][j + 1], F[i + 1][j + 1]
        if M >= e and M >= f:
            return "M"
        return "E" if e >= f else "F"

    ops = []
    i, j, state = t - 1, q - 1, "H"
    while i >= 0 or j >= 0:
        if i < 0:
            ops.append(1); j -= 1; continue
        if j < 0:
            ops.append(2); i -= 1; continue
        if not inband(i, j):
            break
        if state == "H":
            src = came_from(i, j)
            if src == "M":
                ops.append(0); i -= 1; j -= 1
            else:
                state = src
        elif state == "E":            # deletion: consumes target
            ops.append(2)
            opened = H[i][j + 1] - o_del - e_del > E[i][j + 1] - e_del
            if opened:
                state = "H"
            i -= 1
        else:                         # insertion: consumes query
            ops.append(1)
            opened = H[i + 1][j] - o_ins - e_ins > F[i + 1][j] - e_ins
            if opened:
                state = "H"
            j -= 1
    cig = []
    for op in reversed(ops):
        if cig and (cig[-1] & 0xF) == op:
            cig[-1] += 1 << 4
        else:
            cig.append((1 << 4) | op)
    return score, cig

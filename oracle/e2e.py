"""ctypes binding of the end-to-end single-end harness (oracle/e2e_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the swiftbwa_amd
package.  A read batch goes through the oracle's seeding / chaining (phases
1-3), then through extension either on the CPU (`extend_cpu`) or with
regions imported from elsewhere (`set_regions`), then dedup and primary
marking plus the CIGAR job list (`finish`), then CIGARs on the CPU
(`cigar_cpu`) or imported (`set_cigar`), then SAM text (`sam`).
"""
import ctypes as C

import numpy as np

from . import pyoracle as po

_NT4 = np.full(256, 4, np.uint8)
for _i, _c in enumerate(b"ACGT"):
    _NT4[_c] = _i
    _NT4[_c + 32] = _i          # lower case (nst_nt4_table, Base.swift:37-44)

_bound = False


def _lib():
    global _bound
    L = po.lib()
    if _bound:
        return L
    P, I32, I64 = C.c_void_p, C.c_int32, C.c_int64
    sig = {
        "or_e2e_create": (P, [P, I64, C.c_char_p, C.POINTER(po.Scoring)]),
        "or_e2e_free": (None, [P]),
        "or_e2e_set_contigs": (C.c_int, [P, C.c_int, P, P, C.c_char_p]),
        "or_e2e_seq_id": (C.c_int, [P, I64]),
        "or_e2e_fork": (P, [P]),
        "or_e2e_pe_rescue_jobs_raw": (I64, [P, C.c_int]),
        "or_e2e_pe_get_rescue_windows": (None, [P, P, P, P]),
        "or_e2e_pe_set_rescue_windows": (None, [P, P]),
        "or_e2e_sa": (I64, [P, I64]),
        "or_e2e_bwt": (C.c_int, [P, I64]),
        "or_e2e_counts": (None, [P, P]),
        "or_e2e_occ": (I64, [P, I64, C.c_int]),
        "or_e2e_smems": (C.c_int, [P, P, C.c_int, I32, I64, P, C.c_int]),
        "or_e2e_seed": (I64, [P, C.c_int, P, P, P]),
        "or_e2e_n_seeds": (I64, [P]),
        "or_e2e_get_chains": (None, [P, P, P, P, P, P, P]),
        "or_e2e_extend_cpu": (I64, [P]),
        "or_e2e_set_regions": (None, [P, P, P]),
        "or_e2e_n_calls": (I64, [P, P, P]),
        "or_e2e_get_calls": (None, [P] + [P] * 10),
        "or_e2e_finish": (I64, [P]),
        "or_e2e_cigar_sizes": (None, [P, P, P]),
        "or_e2e_get_cigar_jobs": (None, [P] + [P] * 13),
        "or_e2e_cigar_cpu": (None, [P]),
        "or_e2e_set_cigar": (None, [P, I64, P, C.c_int, I32, C.c_char_p, I64]),
        "or_e2e_sam": (I64, [P, P, P, P, I64]),
        "or_e2e_pe_finish": (I64, [P, C.c_int]),
        "or_e2e_pe_dist": (None, [P, P, P]),
        "or_e2e_pe_rescue_jobs": (I64, [P, C.c_int]),
        "or_e2e_pe_rescue_sizes": (None, [P, P, P]),
        "or_e2e_pe_get_rescue_jobs": (None, [P] * 7),
        "or_e2e_pe_rescue_cpu": (None, [P]),
        "or_e2e_pe_set_rescue": (None, [P, P, P]),
        "or_e2e_pe_apply": (I64, [P]),
        "or_e2e_pe_remark": (None, [P]),
        "or_e2e_pe_cigar_plan": (I64, [P]),
        "or_e2e_pe_sam": (I64, [P, P, P, P, I64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _bound = True
    return L


def encode(seq: bytes) -> np.ndarray:
    """ASCII bases -> codes 0..4 (ReadSequence.init(sequence:), nst_nt4)."""
    return _NT4[np.frombuffer(seq, np.uint8)]


def read_fasta(path):
    """First record of a FASTA file: (name, bases 0..4)."""
    name, parts = None, []
    with open(path, "rb") as f:
        for line in f:
            line = line.strip()
            if line.startswith(b">"):
                if name is not None:
                    break
                name = line[1:].split()[0].decode()
            elif line:
                parts.append(line)
    return name, encode(b"".join(parts))


def read_fastq(path):
    """[(name, bases 0..4, quality text bytes)] of a FASTQ file."""
    out = []
    with open(path, "rb") as f:
        lines = [l.rstrip(b"\r\n") for l in f]
    for k in range(0, len(lines) - 3, 4):
        if not lines[k].startswith(b"@"):
            continue
        out.append((lines[k][1:].split()[0].decode(), encode(lines[k + 1]), lines[k + 3]))
    return out


def parse_sam(text):
    """name -> list of records (dict of the SAM columns and typed tags)."""
    recs = {}
    for line in text.splitlines():
        if not line or line.startswith("@"):
            continue
        f = line.split("\t")
        r = {"qname": f[0], "flag": int(f[1]), "rname": f[2], "pos": int(f[3]),
             "mapq": int(f[4]), "cigar": f[5], "seq": f[9], "qual": f[10], "tags": {}}
        for t in f[11:]:
            k, ty, v = t.split(":", 2)
            r["tags"][k] = int(v) if ty == "i" else v
        recs.setdefault(f[0], []).append(r)
    return recs


class E2E:
    """One indexed reference and one read batch at a time.  `contigs`, a list
    of (name, length) tiling the reference, splits it into sequences as an
    index built from a multi-sequence FASTA does; without it the reference is
    one sequence called `name`."""

    def __init__(self, fwd, name="ref", sc=None, contigs=None, _base=None):
        self.L = _lib()
        if _base is not None:                  # fork(): the base's index, own batch
            self.sc, self.fwd, self.gl = _base.sc, _base.fwd, _base.gl
            self.contigs = _base.contigs
            self._base = _base
            self.h = self.L.or_e2e_fork(_base.h)
            return
        self.contigs = [(name, len(fwd))] if not contigs else list(contigs)
        self.sc = sc or po.scoring()
        self.fwd = np.ascontiguousarray(fwd, np.uint8)
        if np.any(self.fwd > 3):
            raise ValueError("reference with ambiguous bases is not supported")
        self.h = self.L.or_e2e_create(self.fwd.ctypes.data, len(self.fwd), name.encode(),
                                      C.byref(self.sc))
        self.gl = len(self.fwd)
        if contigs:
            lens = np.array([n for _, n in contigs], np.int64)
            offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
            names = b"".join(str(nm).encode() + b"\0" for nm, _ in contigs)
            if self.L.or_e2e_set_contigs(self.h, len(contigs), offs.ctypes.data, lens.ctypes.data,
                                         names) != 0:
                self.close()
                raise ValueError("contig lengths must tile the reference")

    def fork(self):
        """A handle over this index with its own batch state (for batches in
        flight on several threads); this handle must stay open meanwhile."""
        return E2E(None, _base=self)

    def seq_id(self, pos):
        return self.L.or_e2e_seq_id(self.h, pos)

    def close(self):
        if getattr(self, "h", None):
            self.L.or_e2e_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # -- index inspection -------------------------------------------------
    def sa(self, i):
        return self.L.or_e2e_sa(self.h, int(i))

    def bwt(self, i):
        return self.L.or_e2e_bwt(self.h, int(i))

    def counts(self):
        c = np.zeros(5, np.int64)
        self.L.or_e2e_counts(self.h, c.ctypes.data)
        return c

    def occ(self, pos, c):
        return self.L.or_e2e_occ(self.h, int(pos), int(c))

    def smems(self, q, min_seed_len=19, min_intv=1):
        q = np.ascontiguousarray(q, np.uint8)
        cap = max(16, 16 * len(q))
        rows = np.zeros((cap, 4), np.int64)
        n = self.L.or_e2e_smems(self.h, q.ctypes.data, len(q), int(min_seed_len), int(min_intv),
                                rows.ctypes.data, cap)
        return [tuple(int(x) for x in rows[i]) for i in range(min(n, cap))]

    # -- pipeline ---------------------------------------------------------
    def seed(self, reads):
        """Phases 1-3 for reads (sequence of base-code arrays).  Returns the
        chains: [(read index, [(rbeg, qbeg, len, score), ...]), ...]."""
        self.reads = [np.ascontiguousarray(r, np.uint8) for r in reads]
        lens = np.array([len(r) for r in self.reads], np.int32)
        off = np.zeros(len(lens), np.int64)
        if len(lens) > 1:
            off[1:] = np.cumsum(lens)[:-1]
        arena = np.concatenate(self.reads) if len(self.reads) and lens.sum() else np.zeros(1, np.uint8)
        arena = np.ascontiguousarray(arena, np.uint8)
        nc = self.L.or_e2e_seed(self.h, len(self.reads), arena.ctypes.data, off.ctypes.data,
                                lens.ctypes.data)
        ns = self.L.or_e2e_n_seeds(self.h)
        cr = np.zeros(max(nc, 1), np.int32)
        co = np.zeros(nc + 1, np.int64)
        rb = np.zeros(max(ns, 1), np.int64)
        qb, sl, ss = (np.zeros(max(ns, 1), np.int32) for _ in range(3))
        self.L.or_e2e_get_chains(self.h, cr.ctypes.data, co.ctypes.data, rb.ctypes.data,
                                 qb.ctypes.data, sl.ctypes.data, ss.ctypes.data)
        self.chains = [(int(cr[c]), [(int(rb[k]), int(qb[k]), int(sl[k]), int(ss[k]))
                                     for k in range(co[c], co[c + 1])]) for c in range(nc)]
        return self.chains

    def extend_cpu(self):
        return self.L.or_e2e_extend_cpu(self.h)

    def calls(self):
        """The BandedSWScalar.align calls of the last extend_cpu as a job dict
        (queries / targets arenas, h0, w, end_bonus) plus the results (n x 6)."""
        qb, tb = C.c_int64(), C.c_int64()
        n = self.L.or_e2e_n_calls(self.h, C.byref(qb), C.byref(tb))
        q = np.zeros(max(qb.value, 1), np.uint8)
        t = np.zeros(max(tb.value, 1), np.uint8)
        qo, to = np.zeros(max(n, 1), np.int64), np.zeros(max(n, 1), np.int64)
        ql, tl, h0, w, eb = (np.zeros(max(n, 1), np.int32) for _ in range(5))
        out = np.zeros((max(n, 1), 6), np.int32)
        self.L.or_e2e_get_calls(self.h, q.ctypes.data, qo.ctypes.data, ql.ctypes.data,
                                t.ctypes.data, to.ctypes.data, tl.ctypes.data, h0.ctypes.data,
                                w.ctypes.data, eb.ctypes.data, out.ctypes.data)
        js = {"n": int(n), "queries": q, "q_off": qo[:n].astype(np.uint32),
              "q_len": ql[:n].astype(np.uint16), "targets": t, "t_off": to[:n].astype(np.uint32),
              "t_len": tl[:n].astype(np.uint16), "h0": h0[:n], "w": w[:n], "end_bonus": eb[:n]}
        return js, out[:n]

    def set_regions(self, per_chain):
        """per_chain: for each chain, a list of dicts with keys qb qe rb re
        score true_score sub sub_n w seed_cov."""
        keys = ("qb", "qe", "rb", "re", "score", "true_score", "sub", "sub_n", "w", "seed_cov")
        n = np.array([len(r) for r in per_chain], np.int32)
        rows = np.array([[g[k] for k in keys] for regs in per_chain for g in regs],
                        np.int64).reshape(-1, 10)
        rows = np.ascontiguousarray(rows if len(rows) else np.zeros((1, 10), np.int64))
        self.L.or_e2e_set_regions(self.h, n.ctypes.data, rows.ctypes.data)

    def finish(self):
        """Phases 4.5-5; returns the CIGAR jobs as a list of dicts."""
        self.n_jobs = self.L.or_e2e_finish(self.h)
        return self._jobs()

    def cigar_cpu(self):
        self.L.or_e2e_cigar_cpu(self.h)

    def _jobs(self):
        qb_, rb_ = C.c_int64(), C.c_int64()
        self.L.or_e2e_cigar_sizes(self.h, C.byref(qb_), C.byref(rb_))
        nj = self.n_jobs
        qs = np.zeros(max(qb_.value, 1), np.uint8)
        rs = np.zeros(max(rb_.value, 1), np.uint8)
        m = max(nj, 1)
        qo, ro, pos = (np.zeros(m, np.int64) for _ in range(3))
        ql, rl, qb, qe, rdl, ts, w = (np.zeros(m, np.int32) for _ in range(7))
        rev = np.zeros(m, np.uint8)
        self.L.or_e2e_get_cigar_jobs(self.h, qs.ctypes.data, qo.ctypes.data, ql.ctypes.data,
                                     rs.ctypes.data, ro.ctypes.data, rl.ctypes.data,
                                     qb.ctypes.data, qe.ctypes.data, rdl.ctypes.data,
                                     rev.ctypes.data, ts.ctypes.data, w.ctypes.data,
                                     pos.ctypes.data)
        self.jobs = [{"qseg": qs[qo[j]:qo[j] + ql[j]].copy(), "rseg": rs[ro[j]:ro[j] + rl[j]].copy(),
                      "qb": int(qb[j]), "qe": int(qe[j]), "read_len": int(rdl[j]),
                      "rev": bool(rev[j]), "true_score": int(ts[j]), "w": int(w[j]),
                      "ref_pos": int(pos[j])} for j in range(nj)]
        return self.jobs

    # -- paired end -------------------------------------------------------
    def seed_pairs(self, reads1, reads2):
        """Phases 1-3 of pairs: read 1 of pair i is read i, read 2 is n + i."""
        assert len(reads1) == len(reads2)
        self.n_pairs = len(reads1)
        return self.seed(list(reads1) + list(reads2))

    def pe_finish(self):
        """Phases 4.5-5 and the insert-size distribution; True when rescue runs."""
        rc = self.L.or_e2e_pe_finish(self.h, self.n_pairs)
        assert rc >= 0
        return bool(rc)

    def pe_dist(self):
        out = np.zeros((4, 8), np.float64)
        pri = C.c_int32()
        self.L.or_e2e_pe_dist(self.h, out.ctypes.data, C.byref(pri))
        return out, pri.value

    def rescue_jobs(self, direction):
        """Local-SW jobs of one rescue direction as a job dict (queries = mate
        sequences, targets = reference windows)."""
        n = self.L.or_e2e_pe_rescue_jobs(self.h, int(direction))
        qb_, tb_ = C.c_int64(), C.c_int64()
        self.L.or_e2e_pe_rescue_sizes(self.h, C.byref(qb_), C.byref(tb_))
        q = np.zeros(max(qb_.value, 1), np.uint8)
        t = np.zeros(max(tb_.value, 1), np.uint8)
        qo, to = np.zeros(max(n, 1), np.int64), np.zeros(max(n, 1), np.int64)
        ql, tl = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32)
        self.L.or_e2e_pe_get_rescue_jobs(self.h, q.ctypes.data, qo.ctypes.data, ql.ctypes.data,
                                         t.ctypes.data, to.ctypes.data, tl.ctypes.data)
        return {"n": int(n), "queries": q, "q_off": qo[:n].astype(np.uint32),
                "q_len": ql[:n].astype(np.uint16), "targets": t,
                "t_off": to[:n].astype(np.uint32), "t_len": tl[:n].astype(np.uint16)}

    def rescue_windows(self, direction):
        """The candidates of one rescue direction before fetchSequence: mate
        queries and raw BWT-space windows [rb, re) with the template's
        sequence id (the caller clamps and fetches); see set_rescue_windows."""
        n = self.L.or_e2e_pe_rescue_jobs_raw(self.h, int(direction))
        qb_, tb_ = C.c_int64(), C.c_int64()
        self.L.or_e2e_pe_rescue_sizes(self.h, C.byref(qb_), C.byref(tb_))
        q = np.zeros(max(qb_.value, 1), np.uint8)
        t = np.zeros(1, np.uint8)
        m = max(n, 1)
        qo, to = np.zeros(m, np.int64), np.zeros(m, np.int64)
        ql, tl = np.zeros(m, np.int32), np.zeros(m, np.int32)
        self.L.or_e2e_pe_get_rescue_jobs(self.h, q.ctypes.data, qo.ctypes.data, ql.ctypes.data,
                                         t.ctypes.data, to.ctypes.data, tl.ctypes.data)
        rb, re = np.zeros(m, np.int64), np.zeros(m, np.int64)
        rid = np.zeros(m, np.int32)
        self.L.or_e2e_pe_get_rescue_windows(self.h, rb.ctypes.data, re.ctypes.data, rid.ctypes.data)
        return {"n": int(n), "queries": q, "q_off": qo[:n].astype(np.uint32),
                "q_len": ql[:n].astype(np.uint16), "rb": rb[:n], "re": re[:n], "rid": rid[:n]}

    def set_rescue_windows(self, clamped):
        clamped = np.ascontiguousarray(clamped, np.int64)
        if len(clamped):
            self.L.or_e2e_pe_set_rescue_windows(self.h, clamped.ctypes.data)

    def rescue_cpu(self):
        self.L.or_e2e_pe_rescue_cpu(self.h)

    def set_rescue(self, out6, ok):
        out6 = np.ascontiguousarray(out6, np.int32).reshape(-1, 6)
        ok = np.ascontiguousarray(ok, np.int32)
        if len(ok):
            self.L.or_e2e_pe_set_rescue(self.h, out6.ctypes.data, ok.ctypes.data)

    def rescue_apply(self):
        return self.L.or_e2e_pe_apply(self.h)

    def pe_cigar_plan(self):
        """remark + generateFilteredCIGARs over both mates; returns the jobs"""
        self.L.or_e2e_pe_remark(self.h)
        self.n_jobs = self.L.or_e2e_pe_cigar_plan(self.h)
        return self._jobs()

    def pe_sam(self, names, quals):
        nb = b"".join(n.encode() + b"\0" for n in names) or b"\0"
        qa = np.frombuffer(b"".join(quals) or b"\0", np.uint8).copy()
        nbuf = C.create_string_buffer(nb, len(nb))
        n = self.L.or_e2e_pe_sam(self.h, nbuf, qa.ctypes.data, None, 0)
        out = C.create_string_buffer(n + 1)
        self.L.or_e2e_pe_sam(self.h, nbuf, qa.ctypes.data, out, n + 1)
        return out.value.decode()

    def run_pe_cpu(self, names, reads1, reads2, quals1, quals2):
        """The whole CPU paired-end path: SAM text."""
        self.seed_pairs(reads1, reads2)
        self.extend_cpu()
        if self.pe_finish():
            for d in (1, 2):
                self.rescue_jobs(d)
                self.rescue_cpu()
                self.rescue_apply()
        self.pe_cigar_plan()
        self.cigar_cpu()
        return self.pe_sam(names, list(quals1) + list(quals2))

    def set_cigar(self, j, cigar, nm, md, pos):
        c = np.ascontiguousarray(cigar if len(cigar) else [0], np.uint32)
        self.L.or_e2e_set_cigar(self.h, int(j), c.ctypes.data, len(cigar), int(nm),
                                md.encode() if isinstance(md, str) else md, int(pos))

    def sam(self, names, quals):
        """SAM records of the batch (read order); quals: Phred+33 bytes per read."""
        nb = b"".join(n.encode() + b"\0" for n in names) or b"\0"
        qa = np.frombuffer(b"".join(quals) or b"\0", np.uint8).copy()
        nbuf = C.create_string_buffer(nb, len(nb))
        n = self.L.or_e2e_sam(self.h, nbuf, qa.ctypes.data, None, 0)
        out = C.create_string_buffer(n + 1)
        self.L.or_e2e_sam(self.h, nbuf, qa.ctypes.data, out, n + 1)
        return out.value.decode()

    def run_cpu(self, names, reads, quals):
        """The whole CPU path: SAM text."""
        self.seed(reads)
        self.extend_cpu()
        self.finish()
        self.cigar_cpu()
        return self.sam(names, quals)

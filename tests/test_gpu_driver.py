"""GPU parity of hipsw_extend_chains (ExtensionAligner.extend for a batch of
chains) against the oracle's sequential restatement.  Every region field is
compared exactly; chains include covered seeds, seeds at read ends, both
strands and decoy seeds."""
import numpy as np
import pytest

from oracle import pyoracle as orc
from swiftbwa_amd import serving
from swiftbwa_amd.metalsw import ExtensionAligner, ScoringParameters

pytestmark = pytest.mark.gpu
KEYS = ("qb", "qe", "rb", "re", "score", "true_score", "sub", "sub_n", "w", "seed_cov")


def oracle_sc(sp):
    return orc.scoring(a=sp.matchScore, b=sp.mismatchPenalty, o_ins=sp.gapOpenPenalty,
                       e_ins=sp.gapExtendPenalty, o_del=sp.gapOpenPenaltyDeletion,
                       e_del=sp.gapExtendPenaltyDeletion, zdrop=sp.zDrop, w=sp.bandWidth,
                       pen_clip5=sp.penClip5, pen_clip3=sp.penClip3,
                       min_seed_len=sp.minSeedLength)


def mutate(rng, s, sub=0.02, indel=0.004):
    out = []
    for b in s:
        u = rng.random()
        if u < indel / 2:
            continue
        if u < indel:
            out.append(int(rng.integers(0, 4)))
        out.append(int((b + rng.integers(1, 4)) % 4) if rng.random() < sub else int(b))
    return out


def find_seeds(read, dbl, lo, hi, k=19):
    """Maximal exact matches of read k-mers inside dbl[lo:hi] (BWT space)."""
    win = dbl[lo:hi].tobytes()
    rb = bytes(read)
    seeds = set()
    for q in range(0, len(read) - k + 1, 9):
        p = win.find(rb[q:q + k])
        if p < 0:
            continue
        r = lo + p
        q0, r0 = q, r
        while q0 > 0 and r0 > 0 and read[q0 - 1] == dbl[r0 - 1]:
            q0 -= 1
            r0 -= 1
        q1, r1 = q + k, r + k
        while q1 < len(read) and r1 < len(dbl) and read[q1] == dbl[r1]:
            q1 += 1
            r1 += 1
        seeds.add((r0, q0, q1 - q0, q1 - q0))
    return sorted(seeds)


def make_chains(seed, n_reads, sp):
    rng = np.random.default_rng(seed)
    gl = 400_000
    ref = serving.make_reference(70 + seed, gl, copies=20, divergence=0.1)
    dbl = np.concatenate([ref, (3 - ref[::-1]).astype(np.uint8)])
    reads, chains = [], []
    for r in range(n_reads):
        rl = int(rng.choice([100, 150, 151]))
        while True:   # windows (<= 10,000 bases) never cross a strand boundary
            pos = int(rng.integers(12000, 2 * gl - 12000 - rl))
            if abs(pos - gl) > 12000:
                break
        read = mutate(rng, dbl[pos:pos + rl])[:rl]
        reads.append(bytes(read))
        seeds = find_seeds(read, dbl, pos - 40, pos + rl + 40)
        if seeds and rng.random() < 0.3:                  # a seed inside another: covered
            s = seeds[int(rng.integers(0, len(seeds)))]
            if s[2] >= 25:
                seeds.append((s[0] + 3, s[1] + 3, s[2] - 6, s[2] - 6))
        if seeds and rng.random() < 0.2:                  # a far decoy scored first
            q0 = int(rng.integers(0, rl - 19))
            ln = int(rng.integers(19, min(40, rl - q0) + 1))
            seeds.append((int(rng.integers(12000, gl - 12000)), q0, ln, 200))
        if seeds:
            order = rng.permutation(len(seeds))
            chains.append((r, [seeds[i] for i in order]))
        if rng.random() < 0.2:                            # decoy chain elsewhere
            q0 = int(rng.integers(0, rl - 19))
            ln = int(rng.integers(19, min(40, rl - q0) + 1))
            rp = int(rng.integers(12000, gl - 12000)) if rng.random() < 0.5 else \
                int(rng.integers(gl + 12000, 2 * gl - 12000))
            chains.append((r, [(rp, q0, ln, ln)]))
    return ref, dbl, reads, chains


@pytest.mark.parametrize("seed", [1, 2])
def test_chains_match_oracle(engine, seed):
    sp = ScoringParameters()
    ref, dbl, reads, chains = make_chains(seed, 600, sp)
    pac = serving.pack_pac(ref)
    st = {}
    got = ExtensionAligner.extendChains(chains, reads, pac, len(ref), sp, engine, stats=st)
    assert len(got) == len(chains)
    osc = oracle_sc(sp)
    n_regions = n_covered = 0
    for c, (r, seeds) in enumerate(chains):
        want = orc.extend_chain(seeds, list(reads[r]), dbl, osc)
        have = [{k: g[k] for k in KEYS} for g in got[c]]
        assert have == want, (c, seeds, have, want)
        n_regions += len(want)
        n_covered += len(seeds) - len(want)
    assert n_regions > 0 and n_covered > 0
    assert st["cells"] > 0


def test_empty_and_read_end_seeds(engine):
    sp = ScoringParameters()
    ref, dbl, reads, _ = make_chains(5, 4, sp)
    pac = serving.pack_pac(ref)
    rl = len(reads[0])
    chains = [(0, [(5000, 0, rl, rl)]),                   # whole read: no extension at all
              (1, [(6000, 0, 19, 19)]),                   # qbeg 0: right side only
              (2, [(7000, len(reads[2]) - 19, 19, 19)]),  # read end: left side only
              (3, [(10, 30, 20, 20)])]                     # left window clipped at 0
    got = ExtensionAligner.extendChains(chains, reads, pac, len(ref), sp, engine)
    osc = oracle_sc(sp)
    for c, (r, seeds) in enumerate(chains):
        want = orc.extend_chain(seeds, list(reads[r]), dbl, osc)
        assert [{k: g[k] for k in KEYS} for g in got[c]] == want, c
    assert ExtensionAligner.extendChains([], reads, pac, len(ref), sp, engine) == []

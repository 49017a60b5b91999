"""Known answers for the oracle's order-dependent state machines, worked out
by hand from the reference's code, each paired with a planted mistake (a
copy of the oracle built with ORC_MUTANT=m) that the answer rejects.

Seed heap and batch cuts (alignment_rmdup, mapper_cpe.c:2124-2426).  A
fabricated index: the read AAACGTCGTCGTAAA (L = 15, one long seed, e = 0, so
all four of its 12-mers are fetched) has two heavy 12-mers, q = 0 (S1) and
q = 3 (S2), 1056 hits each; the other two have none.  Every residual is 0
(the read's flanks are AAA), so every hit passes the 15-mer filter.

  S1[j] = 10j + 1004                          start (p - q) = 10j + 1004
  S2[j] = 10j + 1005 (j < 32)                 start 10j + 1002
          10j + 1007 (32 <= j < 1000)         start 10j + 1004 (= S1's)
          10j + 1008 (j >= 1000)              start 10j + 1005

Phase 1 (:2194-2236) fetches S1[0..31] and S2[0..31] and inserts both in the
seed heap keyed by the RAW last fetched position (:2244-2252): S1 at 1314 is
the root, S2 at 1315 its lone left child.  Every pop of S1 sets the root's key
to S1's next raw last position (:2350), and the sift-down loop
`while (i * 2 < size)` (:1948) never runs at size 2: S1 stays the root until
its 1056 hits are all fetched (32 more chunks, pos_size = 1088), then S2 takes
the root (del_flag, :1940-1942).  S2 is fetched while pos_size + 32 <= 2048
(:2265): 30 chunks, S2[32..991], pos_size = 2048.  The batch is cut (:2356).
  batch 0 = {10j + 1002 : j < 32} U {10j + 1004 : 1 <= j <= 1055}
            (the batch's first fetched position, S1[0], is dropped, :2113)
  batch 1 = S2[992..1055] less its first: {10j + 1004 : 993 <= j <= 999}
            (seven starts verified again, Q9) U {10j + 1005 : 1000 <= j <= 1055}

The 128-block limit (block_size < 256) cannot bind first: a batch's fetches
after the first are 32 positions except each 12-mer's last, so 128 fetches
hold more than 2048 positions with at most 64 selected 12-mers.

Traceback (align_reads, mapper_cpe.c:288-507).  Walking back from the read's
last base, a diagonal step (match :413, substitution :425) is tried before an
insertion (:437) and a deletion (:451), so an indel in a repeat lands at the
repeat's leftmost copy; an insertion on the first step is labelled 'S'
(:393-401) and merged into the M run (:479-495)."""
import numpy as np
import pytest

import oracle_lib as orc
import swmapper_amd as swm

K = 12
READ = b"AAACGTCGTCGTAAA"
CODE = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}


def kmer_code(s):
    return sum(CODE[b] << (2 * i) for i, b in enumerate(s))


def heap_world():
    """The fabricated index above: (hash_count, positions, pos_lseeds, base_2bits)."""
    s1 = [10 * j + 1004 for j in range(1056)]
    s2 = [10 * j + 1005 if j < 32 else (10 * j + 1007 if j < 1000 else 10 * j + 1008) for j in range(1056)]
    c1, c2 = kmer_code(READ[0:12]), kmer_code(READ[3:15])
    hc = np.zeros(4 ** K, np.uint32)
    hc[c1] = hc[c2] = 1056
    pos = np.array(s1 + s2 if c1 < c2 else s2 + s1, np.uint32)
    b2 = np.zeros(2 * 640, np.uint32)            # a genome of 20480 A's
    return hc, pos, np.zeros(pos.size, np.uint16), b2


HEAP_BATCH0 = sorted([10 * j + 1002 for j in range(32)] + [10 * j + 1004 for j in range(1, 1056)])
HEAP_BATCH1 = [10 * j + 1004 for j in range(993, 1000)] + [10 * j + 1005 for j in range(1000, 1056)]


def heap_batches(so):
    hc, pos, ls, b2 = heap_world()
    ref = orc.OracleRef(K, hc, pos, ls, b2, so=so)
    reads = np.zeros((1, 32), np.uint8)
    reads[0, :15] = np.frombuffer(READ, np.uint8)
    _, _, c = ref.trace(reads, np.array([15], np.uint32), 0)
    assert (c["strand"] == 1).sum() == 0          # the reverse strand's 12-mers have no hits
    return [list(map(int, c["cand"][c["batch"] == b])) for b in range(int(c["batch"].max()) + 1)]


def test_seed_heap_and_batch_cut_known_answer():
    got = heap_batches(orc.ORACLE_SO)
    assert len(got) == 2
    assert got[0] == HEAP_BATCH0
    assert got[1] == HEAP_BATCH1
    assert len(set(got[0]) & set(got[1])) == 7     # starts verified again in the later batch


@pytest.mark.parametrize("mutant,mistake", [
    (1, "textbook sift-down: the lone left child is compared, S1 and S2 alternate"),
    (2, "heap keyed by adjusted starts: S2 (1315 - 3) becomes the root first"),
    (3, "batch cut one chunk early: S2[960..991] move to batch 1"),
])
def test_seed_heap_mistakes_are_caught(mutant, mistake):
    got = heap_batches(orc.mutant_so(mutant))
    assert got != [HEAP_BATCH0, HEAP_BATCH1], mistake


LOC = 100_000


def tb_world():
    """A uniform genome with a 5-A run (T before, C after) at LOC+70 and, in
    a second copy, an ACACAC run (G before, T after) at LOC+70."""
    g = swm.synth_genome(200_000, seed=123).copy()
    g[LOC + 69] = ord("T")
    g[LOC + 70:LOC + 75] = ord("A")
    g[LOC + 75] = ord("C")
    g2 = g.copy()
    g2[LOC + 69] = ord("G")
    g2[LOC + 70:LOC + 76] = np.frombuffer(b"ACACAC", np.uint8)
    g2[LOC + 76] = ord("T")
    return g, g2


def tb_cases(g, g2):
    G, G2 = bytes(g), bytes(g2)
    last = next(c for c in b"ACGT" if c not in (G[LOC + 148], G[LOC + 149]))
    return {
        # read: genome[LOC..] with one more A in the run; 1-based start LOC+1
        "homopolymer insertion": (g, G[LOC:LOC + 75] + b"A" + G[LOC + 75:LOC + 149], b"70M1I79M"),
        "homopolymer deletion": (g, G[LOC:LOC + 70] + b"AAAA" + G[LOC + 75:LOC + 151], b"70M1D80M"),
        # a base appended after genome[LOC..LOC+148]: the leftmost best end
        # (Q18) is LOC+148, the first traceback step is an insertion -> 'S'
        "insertion at the last base": (g, G[LOC:LOC + 149] + bytes([last]), b"150M"),
        "dinucleotide insertion": (g2, G2[LOC:LOC + 76] + b"AC" + G2[LOC + 76:LOC + 148], b"70M2I78M"),
        "dinucleotide deletion": (g2, G2[LOC:LOC + 70] + b"ACAC" + G2[LOC + 76:LOC + 152], b"70M2D80M"),
    }


def tb_records(so):
    g, g2 = tb_world()
    refs = {id(g): orc.OracleRef.from_index(orc.OracleIndex(g, K), so=so),
            id(g2): orc.OracleRef.from_index(orc.OracleIndex(g2, K), so=so)}
    out = {}
    for name, (gen, read, _) in tb_cases(g, g2).items():
        reads = np.zeros((1, 256), np.uint8)
        reads[0, :len(read)] = np.frombuffer(read, np.uint8)
        _, recs, _ = refs[id(gen)].trace(reads, np.array([len(read)], np.uint32), 5)
        out[name] = [(int(r["loc"]), bytes(r["cigar"]), int(r["dist"]), int(r["strand"])) for r in recs]
    return out


def test_traceback_ties_known_answers():
    g, g2 = tb_world()
    exp = tb_cases(g, g2)
    got = tb_records(orc.ORACLE_SO)
    for name, (_, read, cigar) in exp.items():
        assert len(read) == 150
        dist = 2 if "dinucleotide" in name else 1
        assert got[name] == [(LOC + 1, cigar, dist, 0)], (name, got[name])


@pytest.mark.parametrize("mutant,case,mistake", [
    (4, "insertion at the last base", "first-step insertion labelled 'I': 149M1I"),
    (5, "homopolymer insertion", "insertion tried before a match: the extra A lands at the run's right end"),
    (5, "dinucleotide insertion", "insertion tried before a match: the extra AC lands at the run's right end"),
])
def test_traceback_mistakes_are_caught(mutant, case, mistake):
    g, g2 = tb_world()
    cigar = tb_cases(g, g2)[case][2]
    got = tb_records(orc.mutant_so(mutant))[case]
    assert got and got[0][1] != cigar, mistake

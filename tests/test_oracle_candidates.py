"""Pins the oracle's candidate batches (alignment_rmdup, mapper_cpe.c:2124-2426).

Closed form for a strand whose selected 12-mers all have <= 32 hits (one
batch, no seed heap): the candidates are the sorted unique values
max(p - (15m+q), 0) over the selected 12-mers (m, q) and their genome hits p
whose 3 flanking bases match the read's long seed (the 15-mer filter), with
the batch's first fetched position left out (the drop at :2113).  The
selected long seeds are the e+1 with the smallest hit totals (ties are
avoided by the case selection).  For strands with more hits the batches
must each be sorted and unique, and their union must be the same set up to
one dropped position per batch.
"""
import numpy as np
import pytest

import oracle_lib as orc
import workloads as wl

CODE = np.full(256, 1, np.int64)
for i, ch in enumerate(b"ACGT"):
    CODE[ch] = i
COMP = bytes.maketrans(b"ACGT", b"TGCA")


def strand_seq(read, strand):
    return read if strand == 0 else read[::-1].translate(COMP)


def stored_code(ix, p):
    w = ix.base_2bits
    return int((w[2 * (p >> 5)] >> (p & 31)) & 1) | (int((w[2 * (p >> 5) + 1] >> (p & 31)) & 1) << 1)


def expected_set(ix, ref_off, rs, e, k=12, any_count=False):
    """Returns (set, n_raw_positions, first_raw) or None when ties make the
    selection order ambiguous or a seed has > 32 hits."""
    L = len(rs)
    M = L // 15
    codes = CODE[np.frombuffer(rs, np.uint8)]
    seeds = {}
    tot = []
    for m in range(M):
        t = 0
        for q in range(4):
            c = int(sum(int(codes[15 * m + q + i]) << (2 * i) for i in range(k)))
            cnt = int(ref_off[c + 1] - ref_off[c])
            seeds[(m, q)] = (c, cnt)
            t += cnt
        tot.append(t)
    nsel = min(e + 1, M)
    order = sorted(range(M), key=lambda m: tot[m])
    if nsel < M and tot[order[nsel - 1]] == tot[order[nsel]]:
        return None                                # selection boundary tied
    pos_tot = sorted(tot[m] for m in order[:nsel] if tot[m] > 0)
    if len(pos_tot) > 1 and pos_tot[0] == pos_tot[1]:
        return None                                # first fetched long seed ambiguous
    sel = [m for m in order[:nsel] if tot[m] > 0]
    if not any_count and any(seeds[(m, q)][1] > 32 for m in sel for q in range(4)):
        return None
    vals, raw, first = [], 0, None
    for m in sel:
        for q in range(4):
            c, cnt = seeds[(m, q)]
            for j in range(cnt):
                p = int(ix.positions[ref_off[c] + j])
                raw += 1
                if first is None and not any_count:
                    first = p
                    continue
                s = max(p - 3, 0)
                ok = all(stored_code(ix, s + 3 - q + t) == codes[15 * m + t] for t in range(q))
                ok = ok and all(stored_code(ix, p + k + t) == codes[15 * m + q + k + t] for t in range(3 - q))
                if ok:
                    vals.append(max(p - (15 * m + q), 0))
    return sorted(set(vals)), raw, first


@pytest.fixture(scope="module")
def small_world():
    g = wl.genome(300_000, seed=77)
    ix = orc.OracleIndex(g, 12)
    ref = orc.OracleRef.from_index(ix)
    return g, ix, ref


def test_single_batch_closed_form(small_world):
    g, ix, ref = small_world
    reads, lens, _, _ = wl.reads(g, 2000, len_min=60, len_max=200, err=(0.0, 0.04), seed=5)
    e = 5
    _, _, cands = ref.trace(reads, lens, e)
    checked = 0
    for r in range(len(lens)):
        read = bytes(reads[r, :lens[r]])
        for strand in (0, 1):
            rs = strand_seq(read, strand)
            exp = expected_set(ix, ref.hash_offset, rs, e)
            if exp is None:
                continue
            sel = cands[(cands["read"] == r) & (cands["strand"] == strand)]
            assert (sel["batch"] == 0).all()
            assert list(sel["cand"]) == exp[0], (r, strand)
            checked += 1
    assert checked > 500


def test_multi_batch_invariants():
    g = wl.genome(400_000, seed=78, kind=2, frac=0.4)       # tandem arrays: long buckets
    ix = orc.OracleIndex(g, 12)
    ref = orc.OracleRef.from_index(ix)
    reads, lens, _, _ = wl.reads(g, 1500, seed=6)
    hits, recs, cands = ref.trace(reads, lens, 5)
    assert cands.size > 0
    multi = 0
    keys = np.unique(cands[["read", "strand"]])
    for key in keys:
        sel = cands[(cands["read"] == key["read"]) & (cands["strand"] == key["strand"])]
        nb = int(sel["batch"].max()) + 1
        multi += nb > 1
        for b in range(nb):
            v = sel["cand"][sel["batch"] == b].astype(np.int64)
            assert (np.diff(v) > 0).all()
            assert list(sel["idx"][sel["batch"] == b]) == list(range(v.size))
        read = bytes(reads[key["read"], :lens[key["read"]]])
        full = expected_set(ix, ref.hash_offset, strand_seq(read, int(key["strand"])), 5, any_count=True)
        if full is not None:
            union = set(sel["cand"].tolist())
            assert union <= set(full[0])
            assert len(set(full[0]) - union) <= nb      # at most one dropped position per batch
    assert multi > 5
    # records only come from verified candidates with err <= e
    assert hits["count"].sum() == recs.size


def test_error_free_reads_map_exactly():
    g = wl.genome(500_000, seed=79)
    rng = np.random.default_rng(1)
    starts = rng.integers(100, g.size - 400, 400)
    strands = rng.integers(0, 2, 400)
    reads, lens = wl.exact_reads(g, starts, 150, strands)
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    hits, _ = ref.map(reads, lens, 5)
    assert (hits["count"] >= 1).all()
    assert (hits["dist"] == 0).all()
    assert (hits["pos"] == starts + 1).all()
    assert (hits["strand"] == strands).all()

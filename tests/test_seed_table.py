"""Layout of the device seed table (CPU, through the library's host export).

The mapping kernel looks up (offset, count) of the k-mers of every long seed
of both strands of a read: four consecutive k-mers c0..c3 per long seed, and
on the reverse strand (for L = 15n) the reverse complements of the same four.
The table stores every k-mer twice -- role 0 keyed by its last k-1 bases,
role 1 by its first k-1 -- so that c_q (role 0) and c_{q+1} (role 1), which
share k-1 bases, and the reverse complements of both, sit in one 128-byte
line (16 entries of 8 bytes).  Pinned here, with codes and reverse
complements computed base by base, independently of the library's bit
tricks: the slot map is injective over (k-mer, role), stays inside the table
(dense when k-1 is odd), and gives every consecutive pair and its reverse
complement one line."""
import numpy as np
import pytest

import swmapper_amd as swm


def rc_code(c, k):
    """Reverse complement of a k-mer code (base i at bits 2i, A C G T = 0..3)."""
    r = 0
    for i in range(k):
        b = (c >> (2 * i)) & 3
        r |= (3 - b) << (2 * (k - 1 - i))
    return r


def slot(c, k, role):
    return swm.lib().swm_seed_slot(c, k, role)


def entries(k):
    return (8 if (k - 1) % 2 else 16) * 4 ** (k - 1)


def next_kmer(c, k, y):
    """The k-mer after c in a read whose next base is y."""
    return (c >> 2) | (y << (2 * (k - 1)))


@pytest.mark.parametrize("k", [2, 3, 4, 5, 6, 7])
def test_every_kmer_injective_and_paired(k):
    n = 4 ** k
    s0 = np.array([slot(c, k, 0) for c in range(n)], dtype=np.uint64)
    s1 = np.array([slot(c, k, 1) for c in range(n)], dtype=np.uint64)
    both = np.concatenate([s0, s1])
    assert both.max() < entries(k)
    assert np.unique(both).size == 2 * n
    if (k - 1) % 2:
        assert np.unique(both // 16).size == 4 ** (k - 1) // 2        # dense: every line used
    for c in range(n):
        for y in range(4):
            d = next_kmer(c, k, y)
            line = s0[c] // 16
            assert s1[d] // 16 == line, (k, c, d)                        # consecutive pair
            assert s1[rc_code(c, k)] // 16 == line, (k, c)               # rc(c) in its role 1
            assert s0[rc_code(d, k)] // 16 == line, (k, d)               # rc(c_{q+1}) in its role 0


@pytest.mark.parametrize("k", [12, 13, 14])
def test_sampled_kmers_paired(k):
    rng = np.random.default_rng(k)
    codes = rng.integers(0, 4 ** k, 3000, dtype=np.uint64)
    seen = {}
    for c in codes.tolist():
        y = c % 4
        d = next_kmer(c, k, y)
        a, b = slot(c, k, 0), slot(d, k, 1)
        ra, rb = slot(rc_code(d, k), k, 0), slot(rc_code(c, k), k, 1)
        for v in (a, b, ra, rb):
            assert v < entries(k)
        assert a // 16 == b // 16 == ra // 16 == rb // 16
        for v, key in ((a, (c, 0)), (b, (d, 1))):
            assert seen.setdefault(v, key) == key, "two k-mers share an entry"


def test_long_seed_takes_two_lines_for_both_strands():
    """A 15-base long seed's four 12-mers and the reverse strand's four
    (reverse complements, looked up in mirrored order) touch two lines."""
    k = 12
    rng = np.random.default_rng(7)
    for _ in range(200):
        bases = rng.integers(0, 4, 15).tolist()
        cq = [sum(bases[q + t] << (2 * t) for t in range(k)) for q in range(4)]
        rq = [rc_code(cq[3 - q], k) for q in range(4)]        # the reverse strand's shifts 0..3
        fwd = {slot(cq[q], k, q & 1) // 16 for q in range(4)}
        rev = {slot(rq[q], k, q & 1) // 16 for q in range(4)}
        assert len(fwd) <= 2 and fwd == rev

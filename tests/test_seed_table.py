"""Layout of the device seed table (CPU, through the library's host export).

The mapping kernel looks up (offset, count) of every k-mer of both strands of
a read.  The reverse strand's k-mers are reverse complements of the forward
strand's, so the table places a k-mer and its reverse complement in one
128-byte line (16 entries of 8 bytes).  Pinned here: the slot map is
injective, stays inside 2 x 4^k entries, and sends c and rc(c) to one line
(the reverse complement computed base by base, independently of the
library's bit tricks)."""
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


def slot(c, k):
    return swm.lib().swm_seed_slot(c, k)


@pytest.mark.parametrize("k", [2, 3, 4, 5, 6, 7])
def test_every_kmer_injective_and_paired(k):
    n = 4 ** k
    s = np.array([slot(c, k) for c in range(n)], dtype=np.uint64)
    assert s.max() < 2 * n
    assert np.unique(s).size == n
    for c in range(n):
        r = rc_code(c, k)
        assert s[c] // 16 == s[r] // 16, (k, c, r)
        if r == c:
            continue
        assert s[c] != s[r]


@pytest.mark.parametrize("k", [12, 13, 14])
def test_sampled_kmers_paired(k):
    rng = np.random.default_rng(k)
    codes = rng.integers(0, 4 ** k, 4000, dtype=np.uint64)
    seen = {}
    for c in codes.tolist():
        a, b = slot(c, k), slot(rc_code(c, k), k)
        assert a < 2 * 4 ** k and b < 2 * 4 ** k
        assert a // 16 == b // 16
        assert seen.setdefault(a, c) == c, "two k-mers share an entry"


def test_palindromes_and_half_used_lines():
    # an even-k palindrome (its own reverse complement) has one entry; the
    # table's lines in use are the representative groups: 4^(k-2) / 2 of
    # them plus the palindromic inner (k-2)-mers, each a 256-byte group
    k = 6
    n = 4 ** k
    groups = {slot(c, k) // 32 for c in range(n)}
    pal_inner = 4 ** ((k - 2) // 2)
    assert len(groups) == (4 ** (k - 2) + pal_inner) // 2
    pal = [c for c in range(n) if rc_code(c, k) == c]
    assert len(pal) == 4 ** (k // 2)

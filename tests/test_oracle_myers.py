"""Pins the oracle's banded Myers (mapper_cpe.c:519-660, compiled form
:668-976) against an explicit banded dynamic programme.

The DP, fixed by reading the recurrence: C[0][j] = 0 for window columns
j in [0, 2e] (free start inside the band), cells outside diagonals
j - i in [0, 2e] are infinite, unit costs; the result is the leftmost end
column loc in [L-1, ref_len-1] with the minimum cost <= e, else
(10000, -1).  A dropped term, a shifted band or a wrong tie rule fails it.
"""
import numpy as np
import pytest

import oracle_lib as orc

INF = 10 ** 9


def band_dp(read, win, L, ref_len, e):
    n = len(win)
    C = [[INF] * (n + 1) for _ in range(L + 1)]
    for j in range(min(n, 2 * e) + 1):
        C[0][j] = 0
    for i in range(1, L + 1):
        for j in range(i, min(n, i + 2 * e) + 1):
            best = C[i - 1][j] + 1
            if j > 0:
                best = min(best, C[i - 1][j - 1] + (read[i - 1] != win[j - 1]), C[i][j - 1] + 1)
            C[i][j] = best
    res = (10000, -1)
    for loc in range(L - 1, ref_len):
        v = C[L][loc + 1]
        if v <= e and v < res[0]:
            res = (v, loc)
    return res


def pack(read_codes, win_codes):
    L = len(read_codes)
    rw = np.zeros(16, np.uint32)
    for t, c in enumerate(read_codes):
        rw[t >> 4] |= np.uint32(c << (2 * (t & 15)))
    nw = (len(win_codes) + 31) // 32 + 2
    lo = np.zeros(nw, np.uint32)
    hi = np.zeros(nw, np.uint32)
    for j, c in enumerate(win_codes):
        lo[j >> 5] |= np.uint32((c & 1) << (j & 31))
        hi[j >> 5] |= np.uint32((c >> 1) << (j & 31))
    return rw, lo, hi


def mutated(rng, win, off, L, p=0.1):
    out, j = [], off
    while len(out) < L:
        u = rng.random()
        if u < p * 0.6:
            out.append(int(rng.integers(4)))
        elif u < p * 0.8:
            j += 1
        elif u < p:
            out.append(win[j] if j < len(win) else 0)
        else:
            out.append(win[j] if j < len(win) else 0)
            j += 1
    return out[:L]


def test_myers_equals_band_dp():
    rng = np.random.default_rng(2024)
    for _ in range(1500):
        e = int(rng.integers(0, 7))
        L = int(rng.integers(3, 60))
        clip = int(rng.integers(0, e + 1)) if rng.random() < 0.3 else 0
        ref_len = L + 2 * e - clip
        win = [int(x) for x in rng.integers(0, 4, L + 2 * e + 40)]
        read = mutated(rng, win, int(rng.integers(0, 2 * e + 1)), L)
        rw, lo, hi = pack(read, win)
        assert orc.myers(rw, lo, hi, L, ref_len, e) == band_dp(read, win, L, ref_len, e)


def test_myers_large_band_and_long_reads():
    rng = np.random.default_rng(7)
    for _ in range(60):
        e = int(rng.integers(8, 16))
        L = int(rng.integers(100, 255))
        win = [int(x) for x in rng.integers(0, 4, L + 2 * e + 40)]
        read = mutated(rng, win, int(rng.integers(0, 2 * e + 1)), L, p=0.05)
        rw, lo, hi = pack(read, win)
        assert orc.myers(rw, lo, hi, L, L + 2 * e, e) == band_dp(read, win, L, L + 2 * e, e)


def test_word_width_32_equals_64():
    """Bits at and above 2e+1 never reach the band, so a 32- and a 64-bit
    band word give the same answers for every e <= 15."""
    rng = np.random.default_rng(11)
    for _ in range(3000):
        e = int(rng.integers(0, 16))
        L = int(rng.integers(15, 255))
        win = [int(x) for x in rng.integers(0, 4, L + 2 * e + 40)]
        read = mutated(rng, win, int(rng.integers(0, 2 * e + 1)), L, p=float(rng.uniform(0, 0.1)))
        rw, lo, hi = pack(read, win)
        rl = L + 2 * e - (int(rng.integers(0, e + 1)) if rng.random() < 0.2 else 0)
        assert orc.myers(rw, lo, hi, L, rl, e, 32) == orc.myers(rw, lo, hi, L, rl, e, 64)


@pytest.mark.parametrize("e", [0, 1, 5, 15])
def test_exact_match_offsets(e):
    rng = np.random.default_rng(e)
    L = 150
    for off in range(2 * e + 1):
        win = [int(x) for x in rng.integers(0, 4, L + 2 * e + 40)]
        read = win[off:off + L]
        rw, lo, hi = pack(read, win)
        err, loc = orc.myers(rw, lo, hi, L, L + 2 * e, e)
        assert err == 0 and loc == L - 1 + off


def test_e0_is_hamming_at_offset_zero():
    rng = np.random.default_rng(3)
    for _ in range(200):
        L = int(rng.integers(20, 200))
        win = [int(x) for x in rng.integers(0, 4, L + 40)]
        read = list(win[:L])
        if rng.random() < 0.5:
            read[int(rng.integers(L))] ^= 1
        rw, lo, hi = pack(read, win)
        err, loc = orc.myers(rw, lo, hi, L, L, 0)
        same = read == win[:L]
        assert (err, loc) == ((0, L - 1) if same else (10000, -1))

"""Known answers for the read encoder (get_seed_from_reads,
mapper_cpe.c:1246-1364, driven seed by seed by choose_seed :1372-1379).

Expected words are the plain 2-bit packing with the exceptions worked out by
hand-executing the reference's statements (Q3):
  * L in {15..180 step 15}: plain packing;
  * L = 195, 210, 225: the tail branch (or a main-loop store right after it)
    re-stores word 11, 12, 13 holding only base 191, 207, 223 in slot 15;
  * L = 160: the tail branch flushes word 9 at base 159, then the trailing
    store at :1361 writes the emptied accumulator over it -> word 9 == 0;
  * L = 240: the same for word 14;
  * L = 241: main-loop stores after the tail branch leave word 14 holding
    only base 239 (slot 15) and word 15 holding base 240;
  * L % 15 in 11..14 (e.g. 56): the tail branch never runs; bases >= 15*L//15
    are never encoded and the partial interleaved word is never stored.
Seeds: seed 4m+q is the 12-mer code at 15m+q, sum(c[15m+q+i] << 2i); a
non-ACGT base contributes i % 4 (its offset in the 12-mer loop) when it
enters through a fresh 12-mer (q = 0) and 3 when it enters as the rolling
base (q > 0); its plane/interleaved code is C (1).
"""
import numpy as np
import pytest

import oracle_lib as orc

CODE = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}


def plain_words(read, n_enc):
    c = [CODE.get(b, 1) for b in read[:n_enc]]
    w = np.zeros(32, np.uint32)
    for t, x in enumerate(c):
        w[16 + (t >> 5) * 2] |= np.uint32((x & 1) << (t & 31))
        w[17 + (t >> 5) * 2] |= np.uint32((x >> 1) << (t & 31))
    full = n_enc // 16
    for j in range(16):
        if j < full or (j == full and n_enc % 16 and n_enc == len(read)):
            v = 0
            for t in range(16 * j, min(16 * j + 16, n_enc)):
                v |= c[t] << (2 * (t - 16 * j))
            w[j] = v
    return w


def rand_read(rng, L):
    return bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, L)])


def seeds_def(read, k=12):
    L = len(read)
    out = []
    for m in range(L // 15):
        for q in range(4):
            p = 15 * m + q
            out.append(sum(CODE[read[p + i]] << (2 * i) for i in range(k)))
    return out


@pytest.mark.parametrize("L", [15, 30, 45, 60, 90, 150, 165, 180])
def test_multiple_of_15_is_plain(L):
    rng = np.random.default_rng(L)
    for _ in range(20):
        r = rand_read(rng, L)
        w, s = orc.encode_read(r)
        assert np.array_equal(w, plain_words(r, L))
        assert list(s) == seeds_def(r)


@pytest.mark.parametrize("L,word,base", [(195, 11, 191), (210, 12, 207), (225, 13, 223)])
def test_single_base_restore(L, word, base):
    rng = np.random.default_rng(L)
    for _ in range(20):
        r = rand_read(rng, L)
        w, s = orc.encode_read(r)
        exp = plain_words(r, L)
        exp[word] = CODE[r[base]] << 30
        assert np.array_equal(w, exp)
        assert list(s) == seeds_def(r)


def test_L160_word9_zero():
    rng = np.random.default_rng(160)
    for _ in range(20):
        r = rand_read(rng, 160)
        w, _ = orc.encode_read(r)
        exp = plain_words(r, 160)
        exp[9] = 0
        assert np.array_equal(w, exp)


def test_L240_word14_zero():
    rng = np.random.default_rng(240)
    for _ in range(20):
        r = rand_read(rng, 240)
        w, _ = orc.encode_read(r)
        exp = plain_words(r, 240)
        exp[14] = 0
        assert np.array_equal(w, exp)


def test_L241_words14_15():
    rng = np.random.default_rng(241)
    for _ in range(20):
        r = rand_read(rng, 241)
        w, _ = orc.encode_read(r)
        exp = plain_words(r, 241)
        exp[14] = CODE[r[239]] << 30
        exp[15] = CODE[r[240]]
        assert np.array_equal(w, exp)


@pytest.mark.parametrize("L", [56, 57, 58, 59, 101, 239])
def test_no_tail_branch(L):
    rng = np.random.default_rng(L)
    for _ in range(20):
        r = rand_read(rng, L)
        w, s = orc.encode_read(r)
        n_enc = 15 * (L // 15)
        exp = plain_words(r, n_enc)
        for j in range(n_enc // 16, 16):
            exp[j] = 0
        assert np.array_equal(w, exp)
        assert list(s) == seeds_def(r)


def test_non_acgt_codes():
    rng = np.random.default_rng(9)
    r = bytearray(rand_read(rng, 45))
    r[5] = ord("N")     # enters seed 0 (q = 0) at loop offset 5 -> contributes 1
    r[12] = ord("N")    # enters seed 1 (q = 1) as the rolling base -> contributes 3
    r[30] = ord("R")    # enters seed 8 (m = 2, q = 0) at offset 0 -> contributes 0
    w, s = orc.encode_read(bytes(r))
    base = [CODE.get(b, None) for b in r]
    def code(m, q):
        v = 0
        for i in range(12):
            p = 15 * m + q + i
            b = base[p]
            if b is None:
                b = {5: 1, 12: 3, 30: 0}[p]
            v |= b << (2 * i)
        return v
    assert list(s) == [code(m, q) for m in range(3) for q in range(4)]
    assert (w[16] >> 5) & 1 == 1 and (w[17] >> 5) & 1 == 0       # plane code C
    assert (w[0] >> 10) & 3 == 1


def test_reverse_complement():
    assert bytes(orc.reverse_read(b"AACGTN")) == b"NACGTT"
    assert bytes(orc.reverse_read(b"acgt")) == b"tgca"

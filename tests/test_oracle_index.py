"""Pins the oracle's index build (serial_index.c, HOST_BUILD) against a
vectorised definition of the same files.

Definition (serial_index.c:146-356,423-426):
  * bases: toupper; A,C,G,T -> 0..3, anything else is N (code 0, mask 1);
  * .base_2bits: low/high bit planes, 32 bases per word pair; .base_3bits:
    N mask; the final flush writes the (empty) accumulators over the last
    word when N is a multiple of 32 (16 for the build-only 2-bit seed words);
  * .hash_count: k-mers at starts p = 0,4,8,.. <= N-k free of N, code
    sum(c[p+i] << 2i) (the rolling seed of get_seed);
  * scatter: for p ascending, the code and N check read the stored words;
    slot = start[code] + writes so far to code; later writes win; slots
    nobody writes hold 0; .hash_offset = start + writes (END offsets) with
    entry 4^k = M;
  * .pos_lseeds: planes of bases p-3..p-1 (0..2 when p < 3) and p+k..p+k+2.
"""
import numpy as np
import pytest

import oracle_lib as orc


def pack_bits(bits, per):
    n = bits.size // per
    w = bits.reshape(n, per).astype(np.uint64) << np.arange(per, dtype=np.uint64)
    return w.sum(axis=1).astype(np.uint32)


def brute(seq, k):
    s = np.frombuffer(seq, dtype=np.uint8) if isinstance(seq, bytes) else seq
    N = s.size
    up = np.where((s >= 97) & (s <= 122), s - 32, s)
    code = np.full(N, -1, np.int64)
    for i, ch in enumerate(b"ACGT"):
        code[up == ch] = i
    valid = code >= 0
    c = np.where(valid, code, 0)
    nw = (N + 31) // 32
    pad = nw * 32 + 64 * 32
    cp = np.zeros(pad, np.int64); cp[:N] = c
    nm = np.zeros(pad, np.int64); nm[:N] = ~valid
    lo = pack_bits(cp & 1, 32); hi = pack_bits(cp >> 1, 32); mk = pack_bits(nm, 32)
    seedw = pack_bits(np.repeat(cp, 2) >> np.tile([0, 1], pad) & 1, 32)   # 16 bases x 2 bits per word
    if N % 32 == 0:
        lo[nw - 1] = hi[nw - 1] = mk[nw - 1] = 0
    if N % 16 == 0:
        seedw[N // 16 - 1] = 0
    # stored view of each base
    sc = (seedw.astype(np.uint64)[:, None] >> (2 * np.arange(16, dtype=np.uint64)) & 3).reshape(-1).astype(np.int64)
    sl = (lo.astype(np.uint64)[:, None] >> np.arange(32, dtype=np.uint64) & 1).reshape(-1).astype(np.int64)
    sh = (hi.astype(np.uint64)[:, None] >> np.arange(32, dtype=np.uint64) & 1).reshape(-1).astype(np.int64)
    sm = (mk.astype(np.uint64)[:, None] >> np.arange(32, dtype=np.uint64) & 1).reshape(-1).astype(np.int64)
    P = np.arange(0, N - k + 1, 4)
    win = P[:, None] + np.arange(k)[None, :]
    shifts = 2 * np.arange(k)
    true_ok = valid[win].all(axis=1)
    true_code = (c[win] << shifts).sum(axis=1)
    nb = 1 << (2 * k)
    count = np.bincount(true_code[true_ok], minlength=nb).astype(np.uint32)
    M = int(count.sum())
    start = np.zeros(nb + 1, np.int64); start[1:] = np.cumsum(count)
    sc_code = (sc[win] << shifts).sum(axis=1)
    sc_ok = (sm[win] == 0).all(axis=1)
    spos = np.maximum(P - 3, 0)
    bef = spos[:, None] + np.arange(3)
    aft = P[:, None] + k + np.arange(3)
    fl = (sl[bef] << np.arange(3)).sum(1) | ((sl[aft] << np.arange(3)).sum(1) << 3)
    fh = (sh[bef] << np.arange(3)).sum(1) | ((sh[aft] << np.arange(3)).sum(1) << 3)
    lseed = fl | (fh << 6)
    writes = np.zeros(nb, np.int64)
    positions = np.zeros(M, np.uint32)
    lseeds = np.zeros(M, np.uint16)
    for p, cc, ok, ls in zip(P, sc_code, sc_ok, lseed):
        if not ok:
            continue
        slot = start[cc] + writes[cc]
        writes[cc] += 1
        if slot < M:
            positions[slot] = p
            lseeds[slot] = ls
    end = np.empty(nb + 1, np.uint32)
    end[:nb] = (start[:nb] + writes).astype(np.uint32)
    end[nb] = M
    b2 = np.empty(2 * nw, np.uint32); b2[0::2] = lo[:nw]; b2[1::2] = hi[:nw]
    return {".base_2bits": b2, ".base_3bits": mk[:nw], ".hash_count": count, ".hash_offset": end,
            ".positions": positions, ".pos_lseeds": lseeds}


def rand_seq(rng, n, n_frac=0.0, lower=0.0):
    s = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, n)].copy()
    if n_frac:
        s[rng.random(n) < n_frac] = ord("N")
    if lower:
        m = rng.random(n) < lower
        s[m] = s[m] + 32
    return s


@pytest.mark.parametrize("n,k,nf", [(1000, 8, 0.0), (1024, 8, 0.0), (1040, 8, 0.01), (4099, 10, 0.002),
                                    (8192, 6, 0.0), (3000, 12, 0.0), (2048, 12, 0.005), (100, 8, 0.0)])
def test_index_matches_definition(n, k, nf):
    rng = np.random.default_rng(n * 31 + k)
    s = rand_seq(rng, n, nf, lower=0.05)
    ix = orc.OracleIndex(s, k)
    exp = brute(s, k)
    for name, a in ix.files().items():
        assert np.array_equal(a, exp[name]), name


def test_final_flush_zeroes_last_word():
    s = np.frombuffer(b"C" * 64, np.uint8)
    ix = orc.OracleIndex(s, 8)
    assert list(ix.base_2bits) == [0xFFFFFFFF, 0, 0, 0]
    s = np.frombuffer(b"C" * 65, np.uint8)
    ix = orc.OracleIndex(s, 8)
    assert list(ix.base_2bits) == [0xFFFFFFFF, 0, 0xFFFFFFFF, 0, 1, 0]


def test_index_invariants_without_tail_effects():
    rng = np.random.default_rng(1)
    s = rand_seq(rng, 50_001, 0.001)
    ix = orc.OracleIndex(s, 10)
    assert ix.hash_count.sum() == ix.positions.size
    start = np.concatenate([[0], np.cumsum(ix.hash_count.astype(np.int64))])
    assert np.array_equal(ix.hash_offset_end[:-1], start[1:])
    for c in np.flatnonzero(ix.hash_count)[:2000]:
        seg = ix.positions[start[c]:start[c + 1]]
        assert (np.diff(seg.astype(np.int64)) > 0).all() and (seg % 4 == 0).all()

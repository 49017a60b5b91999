"""GPU parity: the HIP path against the oracle, bit-exact on every output.

Runs on an MI355X only (marker gpu).  Each test prints the case it runs
before the GPU call so a failing run names it.
"""
import numpy as np
import pytest

import swmapper_amd as swm
import oracle_lib as orc
import workloads as wl

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _compare_index(gi, oi):
    for name, a in oi.files().items():
        b = gi.files()[name]
        assert a.dtype == b.dtype and a.shape == b.shape, name
        bad = np.flatnonzero(a != b)
        assert bad.size == 0, f"{name}: {bad.size} differ, first at {bad[:5]}"


@pytest.mark.parametrize("n,kind,k", [
    (100_000, 0, 12),        # N % 32 == 0: last words zeroed by the final flush
    (100_007, 0, 12),
    (65_552, 0, 12),         # N % 16 == 0, N % 32 != 0
    (200_000, 2, 10),        # tandem repeats, k = 10
    (4_000, 1, 8),
    (1_000_032, 1, 12),      # repeat families, N % 32 == 0
    (4_000_000, 0, 12),      # the small configuration's genome size
    (300_001, 0, 14),        # k = 14: 2k + 12 + start bits = 63, still the 64-bit key sort
])
def test_index_build_matches_oracle(n, kind, k):
    _say("index", n, kind, k)
    g = wl.genome(n, seed=n, kind=kind, frac=0.3)
    _compare_index(swm.build_index(g, k), orc.OracleIndex(g, k))


def test_index_build_n_runs():
    """Long runs of N (many starts dropped from the histogram) and N in the
    last bases."""
    g = wl.genome(400_000, seed=9).copy()
    g[10_000:60_000] = ord("N")
    g[200_000:200_013] = ord("n")
    g[-3:] = ord("N")
    _say("index N runs", g.size)
    _compare_index(swm.build_index(g, 12), orc.OracleIndex(g, 12))


def test_index_build_with_n_and_lowercase():
    g = wl.genome(50_016, seed=3).copy()
    rng = np.random.default_rng(5)
    g[rng.integers(0, g.size, 400)] = ord("N")
    g[1000:1100] = np.char.lower(g[1000:1100].view("S1")).view(np.uint8)
    g[-20:-5] = ord("N")
    _say("index N/lowercase", g.size)
    _compare_index(swm.build_index(g, 12), orc.OracleIndex(g, 12))


def _map_both(g, reads, lens, e, k=12):
    gi = swm.build_index(g, k)
    m = swm.Mapper.from_index(gi)
    got = m.map(reads, lens, e)
    st = m.stats()
    m.close()
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, k))
    exp, nv = ref.map(reads, lens, e)
    return got, exp, st, nv


def _assert_same(got, exp):
    for f in ("count", "pos", "strand", "dist"):
        bad = np.flatnonzero(got[f] != exp[f])
        assert bad.size == 0, f"{f}: {bad.size} reads differ, first {bad[:5]} got {got[bad[:3]]} exp {exp[bad[:3]]}"


def test_map_uniform_150bp_e5():
    g = wl.genome(1_000_000, seed=11)
    r, l, _, _ = wl.reads(g, 20_000, seed=12)
    _say("map uniform 1Mbp 20k x 150 e5")
    got, exp, st, nv = _map_both(g, r, l, 5)
    _assert_same(got, exp)
    assert st["n_verifications"] == nv
    assert (exp["count"] > 0).mean() > 0.5


@pytest.mark.parametrize("e", [0, 3, 7, 8, 15])
def test_map_errors_range(e):
    g = wl.genome(300_000, seed=21)
    r, l, _, _ = wl.reads(g, 4000, len_min=50, len_max=250, err=(0.01, 0.04), seed=22 + e)
    _say("map mixed lengths e", e)
    got, exp, st, nv = _map_both(g, r, l, e)
    _assert_same(got, exp)
    assert st["n_verifications"] == nv


@pytest.mark.parametrize("e", [1, 4, 6, 7])
def test_map_fixed_length_packed_pairs(e):
    """Equal read lengths: every candidate pair runs in one lane's 16-bit halves."""
    g = wl.genome(300_000, seed=23, kind=swm.GENOME_REPEAT_FAMILIES, frac=0.2)
    r, l, _, _ = wl.reads(g, 3000, len_min=150, len_max=150, err=(0.0, 0.05), seed=24 + e)
    _say("map packed pairs e", e)
    got, exp, st, nv = _map_both(g, r, l, e)
    _assert_same(got, exp)
    assert st["n_verifications"] == nv


@pytest.mark.parametrize("L", [15, 29, 50, 56, 59, 60, 100, 149, 150, 151, 160, 200, 214, 240, 241, 250, 254])
def test_map_lengths(L):
    g = wl.genome(200_000, seed=31)
    r, l, _, _ = wl.reads(g, 1500, len_min=L, len_max=L, err=(0.0, 0.03), seed=L)
    _say("map length", L)
    got, exp, st, nv = _map_both(g, r, l, 5)
    _assert_same(got, exp)
    assert st["n_verifications"] == nv


@pytest.mark.parametrize("kind,frac", [(swm.GENOME_REPEAT_FAMILIES, 0.3), (swm.GENOME_TANDEM, 0.3)])
def test_map_repeats(kind, frac):
    g = wl.genome(400_000, seed=41, kind=kind, frac=frac)
    r, l, _, _ = wl.reads(g, 3000, seed=42)
    _say("map repeats", kind)
    got, exp, st, nv = _map_both(g, r, l, 5)
    _assert_same(got, exp)
    assert st["n_verifications"] == nv


def test_map_tandem_mixed_lengths():
    """Many candidates per read and 201 read lengths in one batch: the
    verification order by read length with runs longer than a lane writes
    alone (the wave writes them), next to the heavy compaction."""
    g = wl.genome(400_000, seed=61, kind=swm.GENOME_TANDEM, frac=0.3)
    r, l, _, _ = wl.reads(g, 3000, len_min=50, len_max=250, err=(0.01, 0.04), seed=62)
    _say("map tandem mixed lengths e5")
    got, exp, st, nv = _map_both(g, r, l, 5)
    _assert_same(got, exp)
    assert st["n_verifications"] == nv
    assert nv > 16 * l.size and np.unique(l).size > 100


def test_map_genome_ends_and_n():
    g = wl.genome(64_000, seed=51)
    starts = [0, 1, 2, 3, 4, 5, 7, 31, 32, g.size - 150, g.size - 151, g.size - 160, g.size - 200]
    r, l = wl.exact_reads(g, starts, 150, [0, 1] * 7)
    r2, l2, _, _ = wl.reads(g, 500, seed=52)
    r2[::7, 30] = ord("N")
    r2[::11, 100:104] = ord("N")
    reads = np.vstack([r, r2])
    lens = np.concatenate([l, l2])
    _say("map ends/N")
    got, exp, st, nv = _map_both(g, reads, lens, 5)
    _assert_same(got, exp)


@pytest.mark.parametrize("kind", [swm.GENOME_UNIFORM, swm.GENOME_TANDEM])
def test_seed_kernels_agree(kind):
    """The fixed-k seed kernel (k = 12, the read in registers) and the
    generic one (any k) give the oracle's results on reads of every length
    with non-ACGT bytes at random places -- inside a long seed's first 12
    bases (valued by place) and among its last three (valued 3), on both
    strands."""
    g = wl.genome(300_000, seed=71, kind=kind, frac=0.3)
    r, l, _, _ = wl.reads(g, 4000, len_min=15, len_max=254, err=(0.0, 0.03), seed=72)
    rng = np.random.default_rng(73)
    for i in range(0, l.size, 2):
        for _ in range(int(rng.integers(1, 6))):
            r[i, int(rng.integers(0, l[i]))] = (ord("N"), ord("n"), ord("R"), ord("a"))[int(rng.integers(0, 4))]
    _say("seed kernels agree", kind)
    ix = swm.build_index(g, 12)
    m = swm.Mapper.from_index(ix)
    got = m.map(r, l, 5)
    m.set_option(swm.OPT_SEED_GENERIC, 1)
    got_generic = m.map(r, l, 5)
    m.close()
    exp, _ = orc.OracleRef.from_index(orc.OracleIndex(g, 12)).map(r, l, 5)
    _assert_same(got, exp)
    _assert_same(got_generic, exp)


@pytest.mark.parametrize("k", [10, 13])
def test_map_other_seed_lengths(k):
    """-s other than 12: the generic seed kernel against the oracle."""
    g = wl.genome(200_000, seed=80 + k, kind=swm.GENOME_REPEAT_FAMILIES, frac=0.2)
    # at most 16 long seeds of k + 3 bases per read (the reference's tables, Q24)
    r, l, _, _ = wl.reads(g, 2000, len_min=40, len_max=min(250, 16 * (k + 3)), err=(0.0, 0.03), seed=81 + k)
    _say("map seed length", k)
    got, exp, st, nv = _map_both(g, r, l, 5, k=k)
    _assert_same(got, exp)
    assert st["n_verifications"] == nv


def test_myers_stage_matches_oracle():
    rng = np.random.default_rng(3)
    g = wl.genome(20_000, seed=61)
    ix = orc.OracleIndex(g, 12)
    m = swm.Mapper.from_index(swm.build_index(g, 12))
    n = 4096
    e = 5
    L = rng.integers(40, 255, n).astype(np.uint32)
    ws = rng.integers(0, g.size - 300, n).astype(np.uint32)
    clip = rng.integers(0, e + 1, n).astype(np.uint32)
    rlen = (L + 2 * e - np.where(rng.random(n) < 0.2, clip, 0)).astype(np.uint32)
    words = np.zeros((n, 16), dtype=np.uint32)
    exp_err = np.zeros(n, np.int32)
    exp_loc = np.zeros(n, np.int32)
    b2 = np.zeros(ix.base_2bits.size + 128, np.uint32)
    b2[:ix.base_2bits.size] = ix.base_2bits
    for i in range(n):
        frag = g[ws[i] + rng.integers(0, 2 * e + 1):][:L[i]].copy()
        mut = rng.random(L[i]) < 0.03
        frag[mut] = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, mut.sum())]
        w, _ = orc.encode_read(frag.tobytes(), 12)
        words[i] = w[:16]
        nwin = 12
        lo = np.zeros(nwin, np.uint32)
        hi = np.zeros(nwin, np.uint32)
        for j in range(nwin):
            gpos = int(ws[i]) + 32 * j
            wi, sh = gpos >> 5, gpos & 31
            a0, a1 = int(b2[2 * wi]), int(b2[2 * wi + 2])
            h0, h1 = int(b2[2 * wi + 1]), int(b2[2 * wi + 3])
            lo[j] = ((a0 >> sh) | ((a1 << (32 - sh)) if sh else 0)) & 0xFFFFFFFF
            hi[j] = ((h0 >> sh) | ((h1 << (32 - sh)) if sh else 0)) & 0xFFFFFFFF
        exp_err[i], exp_loc[i] = orc.myers(words[i], lo, hi, int(L[i]), int(rlen[i]), e)
    t = swm.DeviceBuffer.from_array
    dw, dl, dc, dws, drl = t(words), t(L), t(np.arange(n, dtype=np.uint32)), t(ws), t(rlen)
    de = swm.DeviceBuffer(4 * n).zero()
    dlo = swm.DeviceBuffer(4 * n).zero()
    _say("myers stage", n)
    m.myers_device(dw.ptr, dl.ptr, dc.ptr, dws.ptr, drl.ptr, n, e, de.ptr, dlo.ptr)
    got_e, got_l = de.to_array(np.int32), dlo.to_array(np.int32)
    assert np.array_equal(got_e, exp_err)
    assert np.array_equal(got_l, exp_loc)
    assert (exp_err <= e).mean() > 0.3


@pytest.mark.parametrize("n,k", [(400_000, 12), (1_000_032, 12), (200_003, 10)])
def test_index_build_large_genome_path(monkeypatch, n, k):
    """The path genomes past 2^(52-2k) sampled starts take (sort of
    code << 32 | p, counts from the runs, residuals from the planes), forced
    at small sizes; N runs and lowercase included."""
    g = wl.genome(n, seed=n + 1, kind=2, frac=0.3).copy()
    g[5_000:9_000] = ord("N")
    g[-2:] = ord("N")
    g[20_000:20_400] = np.char.lower(g[20_000:20_400].view("S1")).view(np.uint8)
    monkeypatch.setenv("SWM_INDEX_LARGE_PATH", "1")
    _say("index large path", n, k)
    _compare_index(swm.build_index(g, k), orc.OracleIndex(g, k))

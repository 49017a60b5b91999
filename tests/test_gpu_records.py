"""GPU parity of the records path (swm_map_batch_records): every written
record -- read, strand, location, distance, primary flag and CIGAR text --
equals the oracle's, in the oracle's emission order, for light and heavy
reads, every emission kernel and both Myers forms; the per-read results equal
swm_map_batch's.  Then the SAM bytes written from them (library writer and
the swmapper CLI) equal the SAM restatement's.
"""
import gzip
import os
import subprocess
import sys

import numpy as np
import pytest

import swmapper_amd as swm
import oracle_lib as orc
import workloads as wl

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import sam_oracle as so  # noqa: E402

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _assert_records(got, text, exp, what=""):
    assert got.size == exp.size, f"{what}: {got.size} records, oracle {exp.size}"
    for f, g in (("read", "read"), ("strand", "strand"), ("dist", "dist"), ("primary", "primary"), ("pos", "loc")):
        bad = np.flatnonzero(got[f].astype(np.int64) != exp[g].astype(np.int64))
        assert bad.size == 0, f"{what} {f}: {bad.size} differ, first {bad[:5]}: got {got[bad[:3]]} exp {exp[bad[:3]]}"
    cig = swm.cigar_strings(got, text)
    bad = [i for i, c in enumerate(cig) if c != bytes(exp[i]["cigar"])]
    assert not bad, f"{what} cigar: {len(bad)} differ, first {bad[:3]}: {[cig[i] for i in bad[:3]]} " \
                    f"exp {[bytes(exp[i]['cigar']) for i in bad[:3]]}"
    offs = got["cigar_off"].astype(np.int64)
    assert (np.diff(offs) == got["cigar_len"][:-1].astype(np.int64)).all() if got.size > 1 else True
    assert len(text) == (int(offs[-1]) + int(got["cigar_len"][-1]) if got.size else 0)


def _assert_results(got, exp, what=""):
    for f in ("count", "pos", "strand", "dist"):
        bad = np.flatnonzero(got[f] != exp[f])
        assert bad.size == 0, f"{what} {f}: {bad.size} reads differ, first {bad[:5]}"


@pytest.fixture(scope="module")
def work():
    out = {}
    g = wl.genome(300_000, seed=51, kind=swm.GENOME_REPEAT_FAMILIES, frac=0.2)
    r, l, _, _ = wl.reads(g, 2500, len_min=50, len_max=250, err=(0.0, 0.05), seed=52)
    out["repeats"] = (g, r, l)
    g = wl.genome(400_000, seed=41, kind=swm.GENOME_TANDEM, frac=0.3)
    r, l, _, _ = wl.reads(g, 1500, seed=42)
    out["tandem"] = (g, r, l)
    return out


SETTINGS = [
    {},
    {swm.OPT_WORK_QUEUE: 0},
    {swm.OPT_HEAVY_CANDS: 1},
    {swm.OPT_MYERS_PACKED: 0},
]


@pytest.mark.parametrize("name", ["repeats", "tandem"])
@pytest.mark.parametrize("e", [3, 5, 8])
@pytest.mark.parametrize("si", range(len(SETTINGS)))
def test_records_match_oracle(work, name, e, si):
    g, r, l = work[name]
    opts = SETTINGS[si]
    _say("records", name, "e", e, {swm.OPTION_NAMES[k]: v for k, v in opts.items()})
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    hits, orecs, _ = ref.trace(r, l, e)
    m = swm.Mapper.from_index(swm.build_index(g, 12))
    for k, v in opts.items():
        m.set_option(k, v)
    res, recs, text, blocks = m.map_records(r, l, e, blocks=True)
    _assert_results(res, hits, "results")
    _assert_records(recs, text, orecs, f"{name} e={e}")
    ob = orc.block_starts(orecs)
    assert blocks.size == ob.size and (blocks == ob).all(), \
        f"put_cigar blocks: {blocks.size} vs oracle {ob.size}, first difference at " \
        f"{np.flatnonzero(blocks[:min(blocks.size, ob.size)] != ob[:min(blocks.size, ob.size)])[:3]}"
    assert m.stats()["n_records"] == recs.size
    _assert_results(m.map(r, l, e), hits, "map after records")
    m.close()


@pytest.mark.parametrize("si", range(len(SETTINGS)))
def test_reference_buffers_from_device_records(work, si):
    """The reference's ref_pos_buffer / cigar_buffer / unmapped list built
    from the device's records and blocks equal those built from the oracle's,
    byte for byte, and read back by sam_format_2's buffer walk give the
    record-level SAM restatement's lines (tandem reads flush many blocks)."""
    g, r, l = work["tandem"]
    opts = SETTINGS[si]
    _say("reference buffers", {swm.OPTION_NAMES[k]: v for k, v in opts.items()})
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    hits, orecs, _ = ref.trace(r, l, 5)
    m = swm.Mapper.from_index(swm.build_index(g, 12))
    for k, v in opts.items():
        m.set_option(k, v)
    _, recs, text, blocks = m.map_records(r, l, 5, blocks=True)
    m.close()
    pos, cig, unm = swm.ref_blocks(recs, text, blocks, l.size)
    orecs_p = np.zeros(orecs.size, swm.RECORD_DTYPE)
    otext = bytearray()
    for i, x in enumerate(orecs):
        c = bytes(x["cigar"])
        orecs_p[i] = (x["read"], x["loc"], len(otext), x["strand"], x["dist"], x["primary"], len(c))
        otext += c
    epos, ecig, eunm = swm.ref_blocks(orecs_p, bytes(otext), orc.block_starts(orecs), l.size)
    assert pos.shape[0] > len(set(zip(orecs["read"].tolist(), orecs["strand"].tolist())))   # mid-strand flushes
    assert np.array_equal(pos, epos) and np.array_equal(cig, ecig) and np.array_equal(unm, eunm)
    n = l.size
    names = [b"q%d" % i for i in range(n)]
    seqs = [bytes(r[i, :l[i]]) for i in range(n)]
    ref_len, cn, cnl = [0, 400_000], [b"chr1"], [4]
    assert so.sam_format_2_buffers(ref_len, cn, cnl, names, seqs, None, pos, cig, unm) == \
        so.sam_batch(ref_len, cn, cnl, names, seqs, None, orecs)


def test_records_empty_and_unmapped():
    g = wl.genome(100_000, seed=61)
    m = swm.Mapper.from_index(swm.build_index(g, 12))
    _say("records empty")
    res, recs, text = m.map_records(np.zeros((0, 256), np.uint8), np.zeros(0, np.uint32), 5)
    assert res.size == 0 and recs.size == 0 and text == b""
    rng = np.random.default_rng(62)
    junk = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, (50, 256))]
    lens = np.full(50, 120, np.uint32)
    res, recs, text = m.map_records(junk, lens, 2)
    assert (res["count"] == 0).all() and recs.size == 0
    with pytest.raises(swm.SwmError):
        m.map_records(junk, np.full(50, 255, np.uint32), 5)       # L + 2 > stride
    m.close()


def test_sam_from_device_records(work, tmp_path):
    g, r, l = work["repeats"]
    _say("sam from device records")
    prefix = str(tmp_path / "ix")
    with open(prefix + ".ref_info", "wb") as f:
        for name, ln in ((b"chr1", 100_000), (b"chr2", 150_000), (b"chr3", g.size - 250_000)):
            f.write(b"%d\n%d\n%s\n" % (ln, len(name), name))
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    hits, orecs, _ = ref.trace(r, l, 5)
    m = swm.Mapper.from_index(swm.build_index(g, 12))
    res, recs, text = m.map_records(r, l, 5)
    n = l.size
    names = [b"r%d" % i for i in range(n)]
    quals = np.random.default_rng(53).integers(33, 74, r.shape).astype(np.uint8)
    w = swm.SamWriter(str(tmp_path / "o.sam.gz"), prefix)
    w.write(names, r, quals, l, recs, text)
    w.close()
    got = gzip.open(str(tmp_path / "o.sam.gz")).read()
    ref_len, cn, cnl = so.load_ref_info(prefix + ".ref_info")
    exp = so.sam_header(ref_len, cn) + so.sam_batch(ref_len, cn, cnl, names, [bytes(r[i, :l[i]]) for i in range(n)],
                                                    [bytes(quals[i, :l[i]]) for i in range(n)], orecs)
    assert got == exp
    m.close()


def test_cli_writes_reference_sam(tmp_path):
    """swm_index + swmapper on FASTA/FASTQ files: <out>.sam.gz equals the SAM
    restatement over the oracle's records, batch by batch (10,000 reads)."""
    g = wl.genome(150_000, seed=71, kind=swm.GENOME_REPEAT_FAMILIES, frac=0.1)
    r, l, _, _ = wl.reads(g, 12_500, len_min=80, len_max=150, err=(0.0, 0.04), seed=72)
    fa = tmp_path / "ref.fa"
    fa.write_bytes(b">chrQ test contig\n" + b"\n".join(bytes(g[i:i + 70]) for i in range(0, g.size, 70)) + b"\n")
    quals = np.random.default_rng(73).integers(33, 74, r.shape).astype(np.uint8)
    fq = tmp_path / "reads.fq"
    with open(fq, "wb") as f:
        for i in range(l.size):
            f.write(b"@q%d extra\n%s\n+\n%s\n" % (i, bytes(r[i, :l[i]]), bytes(quals[i, :l[i]])))
    here = os.path.dirname(swm.LIB_PATH)
    prefix = str(tmp_path / "ix")
    _say("cli index")
    subprocess.run([os.path.join(here, "swm_index"), "-i", str(fa), "-p", prefix, "-s12"], check=True, timeout=120,
                   stdout=subprocess.DEVNULL)
    _say("cli map")
    subprocess.run([os.path.join(here, "swmapper"), "-p", prefix, f"-i{fq}", f"-o{tmp_path / 'out'}", "-e", "5",
                    "-l256"], check=True, timeout=120, stdout=subprocess.DEVNULL)
    got = gzip.open(str(tmp_path / "out.sam.gz")).read()
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    ref_len, cn, cnl = so.load_ref_info(prefix + ".ref_info")
    exp = so.sam_header(ref_len, cn)
    for lo in range(0, l.size, 10_000):
        hi = min(lo + 10_000, l.size)
        _, orecs, _ = ref.trace(r[lo:hi], l[lo:hi], 5)
        exp += so.sam_batch(ref_len, cn, cnl, [b"q%d" % i for i in range(lo, hi)],
                            [bytes(r[i, :l[i]]) for i in range(lo, hi)],
                            [bytes(quals[i, :l[i]]) for i in range(lo, hi)], orecs)
    assert got == exp


@pytest.mark.parametrize("e", [0, 15])
def test_records_extreme_errors(e):
    """e = 0 (Hamming only) and e = 15 (the widest band, up to 31 runs per
    CIGAR) on mixed lengths."""
    g = wl.genome(200_000, seed=95, kind=swm.GENOME_REPEAT_FAMILIES, frac=0.2)
    r, l, _, _ = wl.reads(g, 1200, len_min=40, len_max=254, err=(0.0, 0.08), seed=96)
    _say("records e", e)
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    hits, orecs, _ = ref.trace(r, l, e)
    m = swm.Mapper.from_index(swm.build_index(g, 12))
    res, recs, text = m.map_records(r, l, e)
    _assert_results(res, hits, f"e={e}")
    _assert_records(recs, text, orecs, f"e={e}")
    m.close()


def test_records_genome_ends_and_n():
    """Reads at both ends of the reference (windows past the end read 0, the
    traceback's start may fall before base 0) and reads with N."""
    g = wl.genome(64_000, seed=51)
    starts = [0, 1, 2, 3, 4, 5, 7, 31, 32, g.size - 150, g.size - 151, g.size - 160, g.size - 200]
    r, l = wl.exact_reads(g, starts, 150, [0, 1] * 7)
    r2, l2, _, _ = wl.reads(g, 500, seed=52)
    r2[::7, 30] = ord("N")
    r2[::11, 100:104] = ord("N")
    reads = np.vstack([r, r2])
    lens = np.concatenate([l, l2])
    _say("records ends/N")
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    hits, orecs, _ = ref.trace(reads, lens, 5)
    m = swm.Mapper.from_index(swm.build_index(g, 12))
    res, recs, text = m.map_records(reads, lens, 5)
    _assert_results(res, hits, "ends")
    _assert_records(recs, text, orecs, "ends")
    m.close()


def test_cli_fasta_reads_write_star_quality(tmp_path):
    """FASTA reads have no qualities: QUAL is '*' on unmapped and primary lines."""
    g = wl.genome(120_000, seed=75)
    r, l, _, _ = wl.reads(g, 700, len_min=60, len_max=150, err=(0.0, 0.05), seed=76)
    r[::9, :] = np.frombuffer(b"ACGT", np.uint8)[np.random.default_rng(77).integers(0, 4, (r[::9].shape))]
    fa = tmp_path / "ref.fa"
    fa.write_bytes(b">c1\n" + bytes(g[:60_000]) + b"\n>c2 second\n" + bytes(g[60_000:]) + b"\n")
    rd = tmp_path / "reads.fa"
    with open(rd, "wb") as f:
        for i in range(l.size):
            f.write(b">f%d\n%s\n" % (i, bytes(r[i, :l[i]])))
    here = os.path.dirname(swm.LIB_PATH)
    prefix = str(tmp_path / "ix")
    _say("cli fasta")
    subprocess.run([os.path.join(here, "swm_index"), "-i", str(fa), "-p", prefix], check=True, timeout=120,
                   stdout=subprocess.DEVNULL)
    subprocess.run([os.path.join(here, "swmapper"), "-p", prefix, f"-i{rd}", f"-o{tmp_path / 'o'}", "-e", "4",
                    "-l200"], check=True, timeout=120, stdout=subprocess.DEVNULL)
    got = gzip.open(str(tmp_path / "o.sam.gz")).read()
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    ref_len, cn, cnl = so.load_ref_info(prefix + ".ref_info")
    assert cn == [b"c1", b"c2"] and ref_len == [0, 60_000, g.size]
    _, orecs, _ = ref.trace(r, l, 4)
    exp = so.sam_header(ref_len, cn) + so.sam_batch(ref_len, cn, cnl, [b"f%d" % i for i in range(l.size)],
                                                    [bytes(r[i, :l[i]]) for i in range(l.size)], None, orecs)
    assert got == exp

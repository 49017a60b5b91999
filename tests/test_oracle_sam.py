"""CIGAR text and SAM output on the CPU.

The oracle's CIGAR strings (align_reads, mapper_cpe.c:277-287, 475-507) are
pinned by reads whose alignment the mathematics fixes: an exact read gives
"<L>M", a read with one inserted or deleted base at a spot with a unique
optimum gives kM1I..M / kM1D..M at the true position, and every record of a
mixed workload consumes exactly its read (M + I = L).  The SAM restatement
(oracle/sam_oracle.py) is pinned by hand-worked lines and by the equivalence
of get_ref_name's search with a plain bisection; the product's writer
(swm_sam.c, host code in libswm.so, no GPU needed) must then give the
oracle's bytes.
"""
import bisect
import gzip
import os
import re
import sys

import numpy as np
import pytest

import oracle_lib as orc
import workloads as wl
import swmapper_amd as swm

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import sam_oracle as so  # noqa: E402


def _trace_one(ref, read, e=5):
    r = np.zeros((1, 256), np.uint8)
    r[0, :len(read)] = np.frombuffer(read, np.uint8)
    _, recs, _ = ref.trace(r, np.array([len(read)], np.uint32), e)
    return [(int(x["loc"]), bytes(x["cigar"]).decode(), int(x["dist"]), int(x["strand"])) for x in recs]


@pytest.fixture(scope="module")
def uniform():
    g = wl.genome(50_000, seed=5)
    return g, bytes(g), orc.OracleRef.from_index(orc.OracleIndex(g, 12))


def test_cigar_exact_and_single_indel(uniform):
    """Length-100 reads (fully encoded by the seed loop): exact, one base
    inserted, one base deleted, the indel base differing from both
    neighbours so the optimum is unique."""
    g, gb, ref = uniform
    rng = np.random.default_rng(3)
    checked = {"exact": 0, "ins": 0, "del": 0}
    L = 100
    for trial in range(60):
        p = int(rng.integers(1000, 40000))
        base = gb[p:p + L + 1]
        k = int(rng.integers(10, 95))
        kind = ("exact", "ins", "del")[trial % 3]
        if kind == "exact":
            read, exp = base[:L], f"{L}M"
        elif kind == "ins":
            x = [c for c in b"ACGT" if c not in (base[k - 1], base[k])][0]
            read, exp = base[:k] + bytes([x]) + base[k:L - 1], f"{k}M1I{L - 1 - k}M"
        else:
            if base[k] in (base[k - 1], base[k + 1]):
                continue
            read, exp = base[:k] + base[k + 1:L + 1], f"{k}M1D{L - k}M"
        got = _trace_one(ref, read)
        want = (p + 1, exp, 0 if kind == "exact" else 1, 0)
        assert want in got, (kind, p, k, got)
        checked[kind] += 1
    assert min(checked.values()) >= 10


def test_cigar_two_edits(uniform):
    """A mismatch merges into the M run (S runs are written as M); an
    insertion and a deletion far apart give kM1I..M1D..M."""
    g, gb, ref = uniform
    rng = np.random.default_rng(11)
    L, n_ok = 100, 0
    for _ in range(40):
        p = int(rng.integers(1000, 40000))
        base = bytearray(gb[p:p + L + 2])
        a, b = int(rng.integers(15, 40)), int(rng.integers(60, 90))
        if base[b] in (base[b - 1], base[b + 1]):
            continue
        # mismatch at a: Hamming-like but through the traceback only if an indel exists; pair it with a deletion at b
        mm = [c for c in b"ACGT" if c != base[a]][0]
        read = bytes(base[:a]) + bytes([mm]) + bytes(base[a + 1:b]) + bytes(base[b + 1:L + 1])
        got = _trace_one(ref, read)
        assert (p + 1, f"{b}M1D{L - b}M", 2, 0) in got, (p, a, b, got)
        x = [c for c in b"ACGT" if c not in (base[a - 1], base[a])][0]
        read = bytes(base[:a]) + bytes([x]) + bytes(base[a:b]) + bytes(base[b + 1:L])
        got = _trace_one(ref, read)
        assert (p + 1, f"{a}M1I{b - a}M1D{L - 1 - b}M", 2, 0) in got, (p, a, b, got)
        n_ok += 1
    assert n_ok >= 20


def _parse(cig):
    return [(int(n), op) for n, op in re.findall(r"(\d+)([MID])", cig)]


def test_cigar_consumes_read_and_text_length():
    """Every record of a mixed workload: M + I = L, "<L>M" for Hamming
    records, len(text) + 2 = the CIGAR buffer bytes the flush rule counts,
    and exactly one primary per (read, strand), its first record."""
    g = wl.genome(200_000, seed=21, kind=swm.GENOME_REPEAT_FAMILIES, frac=0.2)
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    r, l, _, _ = wl.reads(g, 1500, len_min=60, len_max=250, err=(0.01, 0.05), seed=22)
    for e in (3, 8):
        hits, recs, _ = ref.trace(r, l, e)
        assert recs.size > 200
        n_trace = 0
        seen = set()
        for x in recs:
            cig = bytes(x["cigar"]).decode()
            assert "".join(f"{n}{op}" for n, op in _parse(cig)) == cig
            L = int(l[x["read"]])
            ops = _parse(cig)
            assert sum(n for n, op in ops if op in "MI") == L, (cig, L)
            assert len(cig) + 2 == int(x["text_len"])
            if cig != f"{L}M":
                n_trace += 1
            key = (int(x["read"]), int(x["strand"]))
            assert bool(x["primary"]) == (key not in seen)
            seen.add(key)
        assert n_trace > 20


# ---------------------------------------------------------------- SAM oracle


def test_get_ref_name_is_bisection():
    """The reference's search, fast paths included, from any previous
    contig, equals bisect_right(ref_len, loc) - 1 (so a contig's last base,
    loc = ref_len[c + 1], lands on contig c + 1)."""
    rng = np.random.default_rng(5)
    for _ in range(400):
        rn = int(rng.integers(1, 12))
        ref_len = [0]
        for _ in range(rn):
            ref_len.append(ref_len[-1] + int(rng.integers(1, 7)))
        for loc in range(0, ref_len[-1] + 3):
            t = bisect.bisect_right(ref_len, loc) - 1
            for pre in range(0, rn + 1):
                assert so.get_ref_name(ref_len, rn, loc, pre) == t
    assert so.get_ref_name([0, 10, 25], 2, 10, 0) == 1
    assert so.get_ref_name([0, 10, 25], 2, 25, 0) == 2


def test_append_int_and_reverse_read():
    assert so.append_int(0) == b"0"
    assert so.append_int(7) == b"7"
    assert so.append_int(1234567) == b"1234567"
    assert so.append_int(-15) == b"/+"          # '0' + (-1), '0' + (-5)
    assert so.reverse_read(b"ACGTNa", b"ABCDEF") == (b"tNACGT".replace(b"t", b"a"), b"FEDCBA")


def test_sam_lines_by_hand(tmp_path):
    p = tmp_path / "ix.ref_info"
    p.write_bytes(b"10\n4\nchrA\n15\n4\nchrB\n")
    ref_len, names, nl = so.load_ref_info(str(p))
    assert ref_len == [0, 10, 25] and names == [b"chrA", b"chrB"] and nl == [4, 4]
    assert so.sam_header(ref_len, names) == (b"@HD\tVN:1.4\tSO:unsorted\n@SQ\tSN:chrA\tLN:10\n"
                                             b"@SQ\tSN:chrB\tLN:15\n"
                                             b'@PG\tID:XXX\tVN:1.0\tPN:XXX\tCL:"CMDLINE"\n')
    assert so.unmapped_line(b"r1", b"ACGT", b"IIII") == b"r1\t4\t*\t0\t0\t*\t*\t0\t0\tACGT\tIIII\n"
    fwd = so.block_lines(ref_len, names, nl, b"r2", b"ACGTT", b"ABCDE", 0, True,
                         [(3, 1, b"5M"), (10, 0, b"5M"), (12, 12, b"2M1I2M")])
    assert fwd == [b"r2\t0\tchrA\t3\t255\t5M\t*\t0\t0\tACGTT\tABCDE\tNM:i:1\tXM:i:1\n",
                   b"r2\t256\tchrB\t0\t255\t5M\t*\t0\t0\t*\t*\tNM:i:0\tXM:i:0\n",
                   b"r2\t256\tchrB\t2\t255\t2M1I2M\t*\t0\t0\t*\t*\tNM:i:<\tXM:i:<\n"]
    rev = so.block_lines(ref_len, names, nl, b"r3", b"AACGN", b"12345", 1, True, [(25, 2, b"5M")])
    assert rev == [b"r3\t16\t\t0\t255\t5M\t*\t0\t0\tNCGTT\t54321\tNM:i:2\tXM:i:2\n"]
    rev2 = so.block_lines(ref_len, names, nl, b"r3", b"AACGN", b"12345", 1, False, [(4, 2, b"5M")])
    assert rev2 == [b"r3\t272\tchrA\t4\t255\t5M\t*\t0\t0\t*\t*\tNM:i:2\tXM:i:2\n"]


# ---------------------------------------------------------------- product writer


def _product_records(orecs):
    """Oracle records as swm_record_t + CIGAR text."""
    recs = np.zeros(orecs.size, swm.RECORD_DTYPE)
    text = bytearray()
    for i, x in enumerate(orecs):
        c = bytes(x["cigar"])
        recs[i] = (x["read"], x["loc"], len(text), x["strand"], x["dist"], x["primary"], len(c))
        text += c
    return recs, bytes(text)


def _write_ref_info(path, contigs):
    with open(path, "wb") as f:
        for name, ln in contigs:
            f.write(b"%d\n%d\n%s\n" % (ln, len(name), name))


def _expected_sam(prefix, names, seqs, quals, batches):
    ref_len, cn, cnl = so.load_ref_info(prefix + ".ref_info")
    out = so.sam_header(ref_len, cn)
    for (lo, hi, orecs) in batches:
        sub = orecs.copy()
        sub["read"] -= lo
        out += so.sam_batch(ref_len, cn, cnl, names[lo:hi], seqs[lo:hi],
                            quals[lo:hi] if quals is not None else None, sub)
    return out


@pytest.mark.parametrize("gz,with_qual", [(1, True), (0, True), (0, False)])
def test_product_sam_writer_matches_oracle(tmp_path, gz, with_qual):
    g = wl.genome(120_000, seed=31, kind=swm.GENOME_REPEAT_FAMILIES, frac=0.2)
    contigs = [(b"chr1", 40_000), (b"chrM", 1), (b"scaffold_7", 50_000), (b"chrX", 29_999)]
    prefix = str(tmp_path / "ix")
    _write_ref_info(prefix + ".ref_info", contigs)
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    r, l, _, _ = wl.reads(g, 900, len_min=40, len_max=200, err=(0.0, 0.06), seed=32)
    hits, orecs, _ = ref.trace(r, l, 6)
    assert (hits["count"] == 0).any() and (orecs["strand"] == 1).any() and (~orecs["primary"].astype(bool)).any()
    n = l.size
    names = [b"read_%d/1" % i for i in range(n)]
    seqs = [bytes(r[i, :l[i]]) for i in range(n)]
    rng = np.random.default_rng(33)
    qarr = rng.integers(33, 74, r.shape).astype(np.uint8)
    quals = [bytes(qarr[i, :l[i]]) for i in range(n)] if with_qual else None
    # two batches, as the CLI writes them
    cut = 500
    batches = [(0, cut, orecs[orecs["read"] < cut]), (cut, n, orecs[orecs["read"] >= cut])]
    path = str(tmp_path / ("out.sam.gz" if gz else "out.sam"))
    w = swm.SamWriter(path, prefix, gzip_level=gz)
    for lo, hi, sub in batches:
        recs, text = _product_records(sub)
        recs["read"] -= lo
        w.write(names[lo:hi], r[lo:hi], qarr[lo:hi] if with_qual else None, l[lo:hi], recs, text)
    w.close()
    got = gzip.open(path).read() if gz else open(path, "rb").read()
    exp = _expected_sam(prefix, names, seqs, quals, batches)
    assert got == exp


def test_product_sam_writer_contig_edges(tmp_path):
    """Locations at contig ends, the reference end, beyond it (a negative
    location wrapped to u32) and 0, through the product writer."""
    prefix = str(tmp_path / "ix")
    _write_ref_info(prefix + ".ref_info", [(b"a", 10), (b"bb", 15), (b"c", 5)])
    locs = [1, 9, 10, 11, 25, 26, 29, 30, 31, 0, 0xFFFFFFF0, 0x90000000]
    orecs = np.zeros(len(locs), orc.REC_DTYPE)
    for i, loc in enumerate(locs):
        orecs[i]["read"] = i
        orecs[i]["loc"] = loc
        orecs[i]["dist"] = i % 12
        orecs[i]["strand"] = i % 2
        orecs[i]["primary"] = 1
        orecs[i]["cigar"] = b"4M"
    n = len(locs)
    names = [b"q%d" % i for i in range(n)]
    reads = np.zeros((n, 8), np.uint8)
    reads[:, :4] = np.frombuffer(b"ACGT", np.uint8)
    lens = np.full(n, 4, np.uint32)
    recs, text = _product_records(orecs)
    path = str(tmp_path / "e.sam")
    w = swm.SamWriter(path, prefix, gzip_level=0)
    w.write(names, reads, None, lens, recs, text)
    w.close()
    got = open(path, "rb").read()
    exp = _expected_sam(prefix, names, [b"ACGT"] * n, None, [(0, n, orecs)])
    assert got == exp
    lines = got.split(b"\n")
    assert lines[5 + 2].split(b"\t")[2:4] == [b"bb", b"0"]     # loc 10: last base of "a"


def test_sam_writer_rejects_bad_input(tmp_path):
    with pytest.raises(swm.SwmError):
        swm.SamWriter(str(tmp_path / "x.sam"), str(tmp_path / "missing"))
    prefix = str(tmp_path / "ix")
    _write_ref_info(prefix + ".ref_info", [(b"a", 10)])
    w = swm.SamWriter(str(tmp_path / "x.sam"), prefix, gzip_level=0)
    recs = np.zeros(1, swm.RECORD_DTYPE)
    recs["read"] = 5                                              # beyond the batch
    with pytest.raises(swm.SwmError):
        w.write([b"a"], np.zeros((1, 8), np.uint8), None, np.ones(1, np.uint32), recs, b"")
    w.close()


@pytest.mark.parametrize("gz", [0, 1, 6])
def test_product_sam_writer_threads_and_chunks(tmp_path, gz):
    """20k reads (several 8192-line work items): any thread count gives the
    oracle's text; with gzip, the file is one valid multi-member stream."""
    prefix = str(tmp_path / "ix")
    _write_ref_info(prefix + ".ref_info", [(b"c1", 5_000_000), (b"c2", 3_000_000)])
    rng = np.random.default_rng(41)
    n = 20_000
    L = rng.integers(30, 200, n).astype(np.uint32)
    reads = np.frombuffer(b"ACGTN", np.uint8)[rng.integers(0, 5, (n, 256))]
    quals = rng.integers(33, 74, (n, 256)).astype(np.uint8)
    rows = []
    for i in range(n):
        for strand in (0, 1):
            k = int(rng.integers(0, 3)) if rng.random() < 0.7 else 0
            for j in range(k):
                rows.append((i, strand, j == 0, int(rng.integers(0, 8_000_010)), int(rng.integers(0, 6)),
                             b"%dM" % L[i] if rng.random() < 0.5 else b"%dM1I%dM" % (L[i] // 2, L[i] - L[i] // 2 - 1)))
    orecs = np.zeros(len(rows), orc.REC_DTYPE)
    for t, (i, strand, prim, loc, d, cig) in enumerate(rows):
        orecs[t] = (i, strand, d, 0, loc, len(cig) + 2, prim, (0, 0, 0), cig)
    names = [b"read:%d" % i for i in range(n)]
    recs, text = _product_records(orecs)
    exp = _expected_sam(prefix, names, [bytes(reads[i, :L[i]]) for i in range(n)],
                        [bytes(quals[i, :L[i]]) for i in range(n)], [(0, n, orecs)])
    for threads in (1, 3, 16):
        path = str(tmp_path / f"t{threads}.sam")
        w = swm.SamWriter(path, prefix, gzip_level=gz, threads=threads)
        w.write(names, reads, quals, L, recs, text)
        w.close()
        raw = open(path, "rb").read()
        got = gzip.decompress(raw) if gz else raw
        assert got == exp, threads

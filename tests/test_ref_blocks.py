"""The reference's own output buffers on the CPU (swm_records_to_ref_blocks,
host code in libswm.so, no GPU needed).

put_cigar (mapper_cpe.c:1437-1507) hands each flushed block of a (read,
strand) to the host as ref_pos_buffer rows of 258 words (cur_read, count |
main_flag << 31, locations) and cigar_buffer rows of 4096 bytes (distance,
length, text per record), and lists reads with no record by their reverse
strand's cur_read; sam_format_2 (mapper_mpe.c:426-687) reads them back.
Pinned here by a hand-worked batch, by the oracle's blocks read back through
a restatement of sam_format_2's own buffer walk (which must give the SAM the
record-level restatement gives), and by the flush rule every closed block of
the oracle obeys.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as orc
import workloads as wl
import swmapper_amd as swm

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import sam_oracle as so  # noqa: E402


def _records(rows):
    """rows: (read, strand, loc, dist, primary, cigar) -> swm records + text."""
    recs = np.zeros(len(rows), swm.RECORD_DTYPE)
    text = bytearray()
    for i, (rd, st, loc, d, pr, c) in enumerate(rows):
        recs[i] = (rd, loc, len(text), st, d, pr, len(c))
        text += c
    return recs, bytes(text)


def test_hand_worked_batch():
    """Read 0: two forward records (one block, main_flag); read 1: nothing;
    read 2: a reverse record alone in its first block, then one more in a
    second block (no main_flag)."""
    recs, text = _records([(0, 0, 1001, 0, 1, b"150M"), (0, 0, 5003, 2, 0, b"70M1I79M"),
                           (2, 1, 77, 1, 1, b"100M"), (2, 1, 0xFFFFFFFE, 3, 0, b"9M1D91M")])
    blocks = np.array([0, 2, 3], np.uint32)
    pos, cig, unm = swm.ref_blocks(recs, text, blocks, 3)
    assert pos.shape == (3, 258) and cig.shape == (3, 4096)
    assert list(unm) == [3]                                  # read 1's reverse strand: 2 * 1 + 1
    assert list(pos[0, :4]) == [0, 2 | 1 << 31, 1001, 5003] and not pos[0, 4:].any()
    assert list(pos[1, :3]) == [5, 1 | 1 << 31, 77] and not pos[1, 3:].any()
    assert list(pos[2, :3]) == [5, 1, 0xFFFFFFFE] and not pos[2, 3:].any()
    assert bytes(cig[0, :16]) == b"\x00\x04150M\x02\x0870M1I79M" and not cig[0, 16:].any()
    assert bytes(cig[1, :6]) == b"\x01\x04100M" and not cig[1, 6:].any()
    assert bytes(cig[2, :9]) == b"\x03\x079M1D91M" and not cig[2, 9:].any()


def test_empty_and_all_unmapped():
    recs, text = _records([])
    pos, cig, unm = swm.ref_blocks(recs, text, np.zeros(0, np.uint32), 4)
    assert pos.shape == (0, 258) and cig.shape == (0, 4096) and list(unm) == [1, 3, 5, 7]


def test_rejects_blocks_that_do_not_tile():
    recs, text = _records([(0, 0, 10, 0, 1, b"5M"), (0, 1, 20, 0, 1, b"5M")])
    for bad in ([1], [0], [0, 0, 1], [0, 5]):                # not from 0 / across strands / empty / past the end
        with pytest.raises(swm.SwmError):
            swm.ref_blocks(recs, text, np.array(bad, np.uint32), 1)
    with pytest.raises(swm.SwmError):                        # a read beyond the batch
        swm.ref_blocks(recs, text, np.array([0, 1], np.uint32), 0)
    rows = [(0, 0, 10 + i, 0, int(i == 0), b"5M") for i in range(257)]
    recs, text = _records(rows)
    with pytest.raises(swm.SwmError):                        # 257 records in one block
        swm.ref_blocks(recs, text, np.array([0], np.uint32), 1)
    rows = [(0, 0, 10 + i, 0, int(i == 0), b"1M" * 62 + b"1M") for i in range(33)]   # 33 x 128 bytes > 4096
    recs, text = _records(rows)
    with pytest.raises(swm.SwmError):
        swm.ref_blocks(recs, text, np.array([0], np.uint32), 1)


@pytest.fixture(scope="module")
def tandem():
    g = wl.genome(400_000, seed=41, kind=swm.GENOME_TANDEM, frac=0.3)
    r, l, _, _ = wl.reads(g, 400, seed=42)
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    hits, orecs, _ = ref.trace(r, l, 5)
    return g, r, l, hits, orecs


def test_oracle_blocks_follow_the_flush_rule(tandem):
    """Every block the oracle closes before its strand's last one was closed
    by the flush test (count + 8 > 256 or bytes + 256 > 4096, :1790, :1858),
    and no block holds more than 256 records or 4096 bytes.  A plausible
    mistake -- a flush after every batch or only at the strand's end --
    breaks one of these on reads with thousands of records."""
    _, _, _, _, orecs = tandem
    st = orc.block_starts(orecs)
    ends = np.r_[st[1:], orecs.size]
    n_multi = 0
    for b, (a, z) in enumerate(zip(st, ends)):
        seg = orecs[a:z]
        num, nbytes = z - a, int(seg["text_len"].sum())
        assert num <= 256 and nbytes <= 4096
        last = z == orecs.size or (orecs["read"][z], orecs["strand"][z]) != (seg["read"][0], seg["strand"][0])
        if not last:
            n_multi += 1
            assert num + 8 > 256 or nbytes + 256 > 4096, f"block {b} closed early: {num} records, {nbytes} bytes"
        assert bool(seg["primary"][0]) == (a == 0 or (orecs["read"][a - 1], orecs["strand"][a - 1]) !=
                                           (seg["read"][0], seg["strand"][0]))
    assert n_multi >= 5                                      # the fixture does flush mid-strand


def test_buffers_read_back_by_sam_format_2(tandem, tmp_path):
    """The converter's buffers, read the way sam_format_2 reads them, give the
    same SAM lines as the record-level restatement of the same blocks."""
    g, r, l, hits, orecs = tandem
    recs, text = _records([(x["read"], x["strand"], x["loc"], x["dist"], x["primary"], bytes(x["cigar"]))
                           for x in orecs])
    assert (recs["cigar_len"].astype(np.int64) + 2 == orecs["text_len"]).all()
    pos, cig, unm = swm.ref_blocks(recs, text, orc.block_starts(orecs), l.size)
    assert list(unm) == [2 * i + 1 for i in np.flatnonzero(hits["count"] == 0)]
    assert pos.shape[0] == orc.block_starts(orecs).size
    n = l.size
    names = [b"r%d" % i for i in range(n)]
    seqs = [bytes(r[i, :l[i]]) for i in range(n)]
    quals = [bytes([40 + (i % 30)]) * int(l[i]) for i in range(n)]
    ref_len, cn, cnl = [0, 250_000, 400_000], [b"chrA", b"chrB"], [4, 4]
    got = so.sam_format_2_buffers(ref_len, cn, cnl, names, seqs, quals, pos, cig, unm)
    exp = so.sam_batch(ref_len, cn, cnl, names, seqs, quals, orecs)
    assert got == exp

"""SAM restatement of the reference's output stage (mapper_mpe.c).

TEST INFRASTRUCTURE ONLY: the checker the product's SAM writer
(swmapper_amd/csrc/swm_sam.c) is compared against; never imported by the
product.  Pure Python over small batches.

Restates:
  sam_header      mapper_mpe.c:315-341
  reverse_read    mapper_mpe.c:342-366
  get_ref_name    mapper_mpe.c:367-397 (the search loop itself, including its
                  fast paths on the previous contig and the zero entries past
                  ref_len[ref_num], which is a zero-initialised global)
  append_int      mapper_mpe.c:410-425 (C int semantics, negative values included)
  sam_format_2    mapper_mpe.c:426-687 (one put_cigar block per call; the
                  previous-contig hint restarts at 0 for every block, :504, :603)

The reference writes a batch's unmapped reads first, then its blocks, both
in an order set by CPE atomics (put_cigar, mapper_cpe.c:1455-1470).  The
deterministic order checked here: unmapped reads in batch order, then
records by read, forward strand before reverse, emission order within a
strand.  FASTA input has no qualities; the reference then writes whatever its
quality buffer holds, this restatement writes '*'.
"""


def load_ref_info(path):
    """load_ref_info's contig table (mapper_mpe.c:206-234): per contig three
    lines: length, name length, name.  Returns (ref_len cumulative, names,
    name_lens) with ref_len[0] = 0 and one entry per contig end."""
    with open(path, "rb") as f:
        lines = f.read().split(b"\n")
    ref_len, names, name_lens = [0], [], []
    i = 0
    while i + 2 < len(lines) + 0 and lines[i] != b"":
        ref_len.append((ref_len[-1] + int(lines[i])) & 0xFFFFFFFF)
        name_lens.append(int(lines[i + 1]))
        names.append(lines[i + 2])
        i += 3
    return ref_len, names, name_lens


def sam_header(ref_len, names):
    out = [b"@HD\tVN:1.4\tSO:unsorted\n"]
    for i, nm in enumerate(names):
        out.append(b"@SQ\tSN:" + nm + b"\tLN:" + str((ref_len[i + 1] - ref_len[i]) & 0xFFFFFFFF).encode() + b"\n")
    out.append(b'@PG\tID:XXX\tVN:1.0\tPN:XXX\tCL:"CMDLINE"\n')
    return b"".join(out)


def get_ref_name(ref_len, ref_num, ref_loc, pre_ind):
    """The reference's loop, step by step on u32 values (sindex_t)."""
    M = 0xFFFFFFFF

    def rl(j):
        return ref_len[j] if j < len(ref_len) else 0       # zero-initialised global past ref_num

    if rl(pre_ind) <= ref_loc < rl(pre_ind + 1):
        return pre_ind
    if rl(pre_ind + 1) <= ref_loc < rl(pre_ind + 2):
        return pre_ind + 1
    left, right, mid = 0, ref_num, ref_num // 2
    while right > left:
        if ref_loc < rl(mid):
            right = (mid - 1) & M
        elif ref_loc > rl(mid):
            left = mid + 1
        else:
            return mid
        mid = ((left + right) & M) // 2
    if ref_loc < rl(mid):
        return (mid - 1) & M
    return mid


def append_int(n):
    """append_int on a C int: '0' for 0, else the digits of the truncating
    C division and remainder, most significant first."""
    if n == 0:
        return b"0"
    out = []
    while n:
        q = int(n / 10)                   # C truncation toward zero
        r = n - 10 * q
        out.append(48 + r)
        n = q
    return bytes(reversed([c & 0xFF for c in out]))


def _to_int32(u):
    u &= 0xFFFFFFFF
    return u - (1 << 32) if u >= (1 << 31) else u


_COMP = {ord("A"): ord("T"), ord("C"): ord("G"), ord("G"): ord("C"), ord("T"): ord("A")}


def reverse_read(seq, qual):
    rs = bytes(_COMP.get(c, c) for c in reversed(seq))
    rq = bytes(reversed(qual)) if qual is not None else None
    return rs, rq


def unmapped_line(name, seq, qual):
    return (name + b"\t4\t*\t0\t0\t*\t*\t0\t0\t" + seq + b"\t" + (qual if qual is not None else b"*") + b"\n")


def block_lines(ref_len, names, name_lens, name, seq, qual, strand, main_flag, block):
    """sam_format_2's lines for one put_cigar block: block = [(loc, dist, cigar)]."""
    ref_num = len(names)
    nm_pad = [nm.ljust(max(nl, len(nm)), b"\x00") for nm, nl in zip(names, name_lens)] + [b""]
    nl_all = list(name_lens) + [0]
    out = []
    ref_ind = 0
    if strand and main_flag and block:
        seq, qual = reverse_read(seq, qual)
    for j, (loc, dist, cigar) in enumerate(block):
        ref_ind = get_ref_name(ref_len, ref_num, loc, ref_ind)
        first = main_flag and j == 0
        flag = (b"16" if first else b"272") if strand else (b"0" if first else b"256")
        line = [name, b"\t", flag, b"\t", nm_pad[ref_ind][:nl_all[ref_ind]], b"\t",
                append_int(_to_int32(loc - ref_len[ref_ind])), b"\t255\t", cigar, b"\t*\t0\t0\t"]
        if first:
            line += [seq, b"\t", qual if qual is not None else b"*", b"\t"]
        else:
            line += [b"*\t*\t"]
        d = bytes([(48 + dist) & 0xFF])
        line += [b"NM:i:", d, b"\tXM:i:", d, b"\n"]
        out.append(b"".join(line))
    return out


def sam_batch(ref_len, names, name_lens, read_names, seqs, quals, records):
    """One batch: records = oracle records (read, strand, dist, block, loc,
    primary, cigar) in emission order."""
    n = len(read_names)
    has = [False] * n
    for r in records:
        has[int(r["read"])] = True
    out = []
    for i in range(n):
        if not has[i]:
            out.append(unmapped_line(read_names[i], seqs[i], quals[i] if quals is not None else None))
    # regroup into put_cigar blocks: (read, strand, block) runs
    k = 0
    while k < len(records):
        r0 = records[k]
        key = (int(r0["read"]), int(r0["strand"]), int(r0["block"]))
        blk = []
        while k < len(records) and (int(records[k]["read"]), int(records[k]["strand"]),
                                    int(records[k]["block"])) == key:
            blk.append((int(records[k]["loc"]), int(records[k]["dist"]), bytes(records[k]["cigar"])))
            k += 1
        i = key[0]
        out += block_lines(ref_len, names, name_lens, read_names[i], seqs[i],
                           quals[i] if quals is not None else None, key[1], bool(r0["primary"]), blk)
    return b"".join(out)


def _schar(b):
    return b - 256 if b >= 128 else b


def sam_format_2_buffers(ref_len, names, name_lens, read_names, seqs, quals, ref_pos, cigar_buffer, unmapped):
    """sam_format_2 (mapper_mpe.c:426-687) reading the reference's own buffers:
    unmapped = cigar_num2[0 .. cigar_unmap_size_pre) (read = entry / 2, :445-482);
    ref_pos = ref_pos_buffer as rows of LOCATIONS = 258 words, cigar_buffer as
    rows of CIGARS_PER_BLOCK * CIGAR_AVG_LEN = 4096 bytes (:486-492): per block
    cur_read, cigar_num (main_flag = bit 31, count = low 28 bits, :494-495),
    locations; per record a distance char, a length char, the text (:536-539)."""
    out = []
    for u in unmapped:
        i = int(u) // 2
        out.append(unmapped_line(read_names[i], seqs[i], quals[i] if quals is not None else None))
    for b in range(len(ref_pos)):
        row = ref_pos[b]
        cur_read = int(row[0])
        i, strand = cur_read // 2, cur_read & 1
        num = int(row[1])
        main_flag = num >> 31
        num &= 0xfffffff
        cig = bytes(cigar_buffer[b])
        at = 0
        blk = []
        for j in range(num):
            dist = _schar(cig[at])
            clen = _schar(cig[at + 1])
            blk.append((int(row[2 + j]), dist, cig[at + 2:at + 2 + clen]))
            at += 2 + clen
        out += block_lines(ref_len, names, name_lens, read_names[i], seqs[i],
                           quals[i] if quals is not None else None, strand, bool(main_flag), blk)
    return b"".join(out)

"""The FASTA/FASTQ reader against kseq's parse rules (kseq_read,
kseq.h:197-237; ks_getuntil2, :99-147), as known answers worked out from
those lines: names end at any isspace() character, one trailing CR per line
is dropped once the sequence is longer than one character, a header at the
end of the file is an empty record, quality lines are taken whole."""
import os

import swmapper_amd as swm


def parse(tmp_path, data, name="x.fa"):
    p = os.path.join(tmp_path, name)
    with open(p, "wb") as f:
        f.write(data)
    return swm.read_fastx(p)


FASTA = b">chr1 the first\nACGTN\nacg\n\n>chr2\nGG\n"
FASTQ = b"@r1 x\nACGT\n+\nIIII\n@r2\nAC\n+r2\nI#\n"


def test_fasta_lf_and_crlf_are_the_same(tmp_path):
    lf, rc = parse(tmp_path, FASTA)
    assert rc == -1
    assert lf == [(b"chr1", b"ACGTNacg", b""), (b"chr2", b"GG", b"")]
    crlf, rc = parse(tmp_path, FASTA.replace(b"\n", b"\r\n"))
    assert rc == -1 and crlf == lf


def test_fastq_lf_and_crlf_are_the_same(tmp_path):
    lf, rc = parse(tmp_path, FASTQ, "x.fq")
    assert rc == -1
    assert lf == [(b"r1", b"ACGT", b"IIII"), (b"r2", b"AC", b"I#")]
    crlf, rc = parse(tmp_path, FASTQ.replace(b"\n", b"\r\n"), "x.fq")
    assert rc == -1 and crlf == lf


def test_name_ends_at_any_space_character(tmp_path):
    for sep in (b" ", b"\t", b"\v", b"\f", b"\r"):
        recs, _ = parse(tmp_path, b">na" + sep + b"me\nAC\n")
        assert recs == [(b"na", b"AC", b"")], sep


def test_header_at_end_of_file_is_an_empty_record(tmp_path):
    recs, rc = parse(tmp_path, b">a\nAC\n>b")
    assert rc == -1 and recs == [(b"a", b"AC", b""), (b"b", b"", b"")]
    recs, rc = parse(tmp_path, b">a\nAC\n>")
    assert recs == [(b"a", b"AC", b"")]           # nothing after '>': end of file


def test_cr_only_first_line_keeps_its_cr(tmp_path):
    # the first line appends '\r' alone (length 1: not stripped); the next
    # line "AC\r" makes "\rAC\r" and loses its trailing CR
    recs, _ = parse(tmp_path, b">y\r\n\r\nAC\r\n")
    assert recs == [(b"y", b"\rAC", b"")]


def test_quality_lines_are_taken_whole(tmp_path):
    recs, rc = parse(tmp_path, b"@q\nACGT\n+\nII\nII\n", "x.fq")
    assert rc == -1 and recs == [(b"q", b"ACGT", b"IIII")]
    recs, rc = parse(tmp_path, b"@q\nACGT\n+\nIIIII\n", "x.fq")   # one too many
    assert rc == -2
    recs, rc = parse(tmp_path, b"@q\nACGT\n+\nII\n", "x.fq")      # too few at the end
    assert rc == -2
    recs, rc = parse(tmp_path, b"@q\nACGT\n+", "x.fq")            # no quality line
    assert rc == -2


def test_text_before_the_first_header_is_skipped(tmp_path):
    recs, _ = parse(tmp_path, b"junk\n\n>a b\nA C\n")
    assert recs == [(b"a", b"A C", b"")]

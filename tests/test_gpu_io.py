"""Index files from a FASTA through the kseq-compatible reader: the CRLF
form of a multi-contig FASTA gives the same files, byte for byte, as its LF
form, and both equal the oracle's build of the concatenated contigs."""
import os

import numpy as np
import pytest

import swmapper_amd as swm
import oracle_lib as orc

pytestmark = pytest.mark.gpu


def write_fasta(path, contigs, eol):
    with open(path, "wb") as f:
        for name, seq in contigs:
            f.write(b">" + name + b" some description" + eol)
            for i in range(0, len(seq), 60):
                f.write(seq[i:i + 60] + eol)


def test_crlf_fasta_index_identical(tmp_path):
    g = swm.synth_genome(300_000, seed=17)
    contigs = [(b"chrA", bytes(g[:120_001])), (b"chrB", bytes(g[120_001:300_000]))]
    files = {}
    for tag, eol in (("lf", b"\n"), ("crlf", b"\r\n")):
        fa = os.path.join(tmp_path, tag + ".fa")
        write_fasta(fa, contigs, eol)
        prefix = os.path.join(tmp_path, tag)
        print("[case] index from", tag, flush=True)
        swm.index_build_files(fa, prefix, 12)
        files[tag] = {ext: open(prefix + ext, "rb").read()
                      for ext in (".ref_info", ".base_2bits", ".base_3bits", ".hash_count", ".hash_offset",
                                  ".positions", ".pos_lseeds")}
    assert files["lf"] == files["crlf"]
    assert files["lf"][".ref_info"] == b"120001\n4\nchrA\n179999\n4\nchrB\n"
    oi = orc.OracleIndex(g, 12)
    for ext, arr in oi.files().items():
        assert files["lf"][ext] == arr.tobytes(), ext

"""CPU checks of the product library: it loads, exports every entry point
declared in include/*.h, validates arguments before touching the GPU, and
writes the reference's index file layout."""
import ctypes as C
import os
import re

import numpy as np

import swmapper_amd as swm
import oracle_lib as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(swm_\w+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    L = swm.lib()
    names = declared("swm.h") + declared("swm_synth.h")
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_argument_validation_without_gpu():
    L = swm.lib()
    out = np.zeros(4, swm.RESULT_DTYPE)
    reads = np.zeros((4, 256), np.uint8)
    lens = np.full(4, 150, np.uint32)
    assert L.swm_map_batch(None, reads.ctypes.data, 256, lens.ctypes.data, 4, 5, out.ctypes.data) == -1
    assert L.swm_map_batch_device_noalloc(None, None, 256, None, 4, 5, None) == -1
    h = C.c_void_p()
    assert L.swm_index_load(b"/nonexistent/prefix", 12, C.byref(h)) == -2
    assert L.swm_index_load(b"/nonexistent/prefix", 1, C.byref(h)) == -1
    assert L.swm_ctx_create(1, None, None, None, 0, None, 0, C.byref(h)) == -1
    ih = swm.IndexHost()
    assert L.swm_index_build_mem(None, 10, 12, C.byref(ih), None) == -1
    assert L.swm_index_build(b"/nonexistent.fa", b"/tmp/x", 12) == -2
    v = C.c_int64(0)
    assert L.swm_ctx_set_option(None, swm.OPT_GRAPH, 1) == -1
    assert L.swm_ctx_get_option(None, swm.OPT_GRAPH, C.byref(v)) == -1
    assert L.swm_autotune(None, reads.ctypes.data, 256, lens.ctypes.data, 4, 5, 1, None) == -1
    cfg = swm.ServerConfig(4, 5, 256, 0, 0, 0, 1, 0)
    assert L.swm_server_create(None, C.byref(cfg), C.byref(h)) == -1
    t = C.c_uint64(0)
    assert L.swm_server_submit(None, reads.ctypes.data, lens.ctypes.data, 4, out.ctypes.data, C.byref(t)) == -1
    assert L.swm_server_wait(None, 1, None) == -1
    L.swm_server_destroy(None)


def test_index_files_layout(tmp_path):
    g = swm.synth_genome(10_000, seed=4)
    oi = orc.OracleIndex(g, 8)
    ix = swm.Index(8, oi.n_bases, oi.base_2bits, oi.base_3bits, oi.hash_count, oi.hash_offset_end,
                   oi.positions, oi.pos_lseeds)
    prefix = str(tmp_path / "ref")
    ix.write(prefix, "chrT")
    for ext, arr in ix.files().items():
        raw = open(prefix + ext, "rb").read()
        assert raw == arr.tobytes(), ext
    assert open(prefix + ".ref_info").read() == "10000\n4\nchrT\n"


def test_mis_sized_index_rejected(tmp_path):
    g = swm.synth_genome(10_000, seed=4)
    oi = orc.OracleIndex(g, 8)
    ix = swm.Index(8, oi.n_bases, oi.base_2bits, oi.base_3bits, oi.hash_count, oi.hash_offset_end,
                   oi.positions, oi.pos_lseeds)
    prefix = str(tmp_path / "ref")
    ix.write(prefix)
    h = C.c_void_p()
    # wrong -s for these files: .hash_count has 4^8 entries, not 4^9
    assert swm.lib().swm_index_load(prefix.encode(), 9, C.byref(h)) == -2
    with open(prefix + ".positions", "ab") as f:
        f.write(b"\0\0\0\0")
    assert swm.lib().swm_index_load(prefix.encode(), 8, C.byref(h)) == -2


def test_synthetic_reads_follow_the_model():
    g = swm.synth_genome(200_000, seed=9)
    r, l, tp, ts = swm.synth_reads(g, 2000, len_min=50, len_max=250, err_min=0.0, err_max=0.0, seed=3)
    assert l.min() >= 50 and l.max() <= 250
    comp = bytes.maketrans(b"ACGT", b"TGCA")
    for i in range(200):
        frag = g[tp[i]:tp[i] + l[i]].tobytes()
        read = r[i, :l[i]].tobytes()
        assert read == (frag[::-1].translate(comp) if ts[i] else frag)
    assert 0.4 < ts.mean() < 0.6

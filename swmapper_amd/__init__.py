"""swmapper_amd -- MI355X-native mapping hot path of SWMapper behind a C ABI.

Python is a thin host mirror of ``include/swm.h`` (ctypes); all compute runs
in ``libswm.so`` (HIP kernels for gfx950).  Loading fails loudly when the
library has not been built: there is no CPU fallback.

Mirrors of the reference's interface:
  * ``build_index(seq, k)``           -- serial_index.c construct_hash_index (:359-445)
  * ``index_build_files(fa, p, k)``   -- serial_index.c main (:818-865)
  * ``Mapper.from_files(prefix, k)``  -- load_ref_info (mapper_mpe.c:198-314)
  * ``Mapper.map(reads, lens, e)``    -- athread_spawn(find_positions) + join
                                         (mapper_mpe.c:972-1036)
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libswm.so")

SWM_OK = 0
_ERRORS = {-1: "invalid argument", -2: "index/sequence file error", -3: "out of memory",
           -4: "HIP device error", -5: "busy", -6: "batch exceeds the no-allocation scratch"}
SWM_ENOSPC = -6
NOALLOC_MAX_READS = 8192

RESULT_DTYPE = np.dtype([("pos", "<u4"), ("strand", "u1"), ("dist", "u1"), ("pad", "<u2"), ("count", "<u4")])
# swm_record_t: one written record (include/swm.h)
RECORD_DTYPE = np.dtype([("read", "<u4"), ("pos", "<u4"), ("cigar_off", "<u4"), ("strand", "u1"), ("dist", "u1"),
                         ("primary", "u1"), ("cigar_len", "u1")])
UNMAPPED_POS = 0xFFFFFFFF

GENOME_UNIFORM, GENOME_REPEAT_FAMILIES, GENOME_TANDEM = 0, 1, 2


class SwmError(RuntimeError):
    def __init__(self, fn, rc):
        super().__init__(f"{fn} failed: {_ERRORS.get(rc, 'status')} ({rc})")
        self.rc = rc


class IndexHost(C.Structure):
    _fields_ = [("k", C.c_uint32), ("n_bases", C.c_uint64), ("n_words", C.c_uint64),
                ("base_2bits", C.POINTER(C.c_uint32)), ("base_3bits", C.POINTER(C.c_uint32)),
                ("hash_count", C.POINTER(C.c_uint32)), ("hash_offset_end", C.POINTER(C.c_uint32)),
                ("positions", C.POINTER(C.c_uint32)), ("pos_lseeds", C.POINTER(C.c_uint16)),
                ("n_pos", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("n_reads", C.c_uint64), ("n_verifications", C.c_uint64), ("n_band_steps", C.c_uint64),
                ("n_records", C.c_uint64), ("n_positions", C.c_uint64),
                ("ms_seed", C.c_double), ("ms_plan", C.c_double), ("ms_compact", C.c_double),
                ("ms_myers", C.c_double), ("ms_emit", C.c_double), ("ms_total", C.c_double),
                ("ms_host_enqueue", C.c_double), ("ms_host_wait", C.c_double),
                ("cand_hist", C.c_uint64 * 24), ("max_cands", C.c_uint64),
                ("n_heavy_items", C.c_uint64), ("n_large_batches", C.c_uint64), ("n_heavy_reads", C.c_uint64),
                ("n_giant_items", C.c_uint64), ("n_giant_reads", C.c_uint64), ("n_irregular_items", C.c_uint64),
                ("n_graph_captures", C.c_uint64), ("n_light_positions", C.c_uint64), ("ms_plan_light", C.c_double)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["cand_hist"] = list(self.cand_hist)
        return d


class SeqRec(C.Structure):
    _fields_ = [("name", C.c_void_p), ("seq", C.c_void_p), ("qual", C.c_void_p),
                ("name_len", C.c_size_t), ("seq_len", C.c_size_t), ("qual_len", C.c_size_t),
                ("name_cap", C.c_size_t), ("seq_cap", C.c_size_t), ("qual_cap", C.c_size_t)]


class Records(C.Structure):
    _fields_ = [("n", C.c_uint64), ("rec", C.c_void_p), ("text_bytes", C.c_uint64), ("cigar", C.c_void_p),
                ("n_blocks", C.c_uint64), ("block", C.c_void_p)]


class RefBlocks(C.Structure):
    _fields_ = [("n_blocks", C.c_uint64), ("ref_pos", C.c_void_p), ("cigar", C.c_void_p),
                ("n_unmapped", C.c_uint64), ("unmapped", C.c_void_p)]


REF_MAX_CIGAR = 256                      # SWM_REF_MAX_CIGAR
REF_POS_STRIDE = REF_MAX_CIGAR + 2       # u32 words per block of ref_pos_buffer
REF_TEXT_STRIDE = REF_MAX_CIGAR * 16     # bytes per block of cigar_buffer


class ServerConfig(C.Structure):
    _fields_ = [("n_slots", C.c_uint32), ("e", C.c_uint32), ("stride", C.c_uint64), ("hbm_budget", C.c_uint64),
                ("max_batch", C.c_uint64), ("max_launch", C.c_uint64), ("graph", C.c_uint32), ("pad", C.c_uint32)]


class BatchInfo(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("n_reads", C.c_uint64), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("status", C.c_int32), ("slot", C.c_uint32),
                ("est_bytes", C.c_uint64), ("launch_reads", C.c_uint64), ("launch_batches", C.c_uint32),
                ("pad", C.c_uint32), ("stats", Stats)]


# forms and tunables (swm_ctx_set_option)
OPT_WORK_QUEUE, OPT_GRAPH, OPT_HEAVY_CANDS, OPT_MYERS_BLOCK, OPT_PLAN_HEAVY_WAVES, OPT_EMIT_HEAVY_WAVES = range(1, 7)
OPT_HISTOGRAM = 7
OPT_MYERS_PACKED = 8
OPT_GIANT_POS, OPT_GIANT_CANDS = 9, 10
OPT_SEED_GENERIC = 11
OPT_PLAN_OVERLAP, OPT_PLAN_LIGHT_WAVES = 12, 13
OPTION_NAMES = {OPT_WORK_QUEUE: "work_queue", OPT_GRAPH: "graph", OPT_HEAVY_CANDS: "heavy_cands",
                OPT_MYERS_BLOCK: "myers_block", OPT_PLAN_HEAVY_WAVES: "plan_heavy_waves",
                OPT_EMIT_HEAVY_WAVES: "emit_heavy_waves", OPT_HISTOGRAM: "histogram",
                OPT_MYERS_PACKED: "myers_packed", OPT_GIANT_POS: "giant_pos", OPT_GIANT_CANDS: "giant_cands",
                OPT_SEED_GENERIC: "seed_generic", OPT_PLAN_OVERLAP: "plan_overlap",
                OPT_PLAN_LIGHT_WAVES: "plan_light_waves"}
N_OPTIONS = 13


_lib_handle = None


def lib():
    """The product library; raises if it has not been built."""
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m swmapper_amd.build` (no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32
        sig = {
            "swm_index_build": ([C.c_char_p, C.c_char_p, u32], C.c_int),
            "swm_index_build_mem": ([vp, u64, u32, C.POINTER(IndexHost), C.POINTER(C.c_double)], C.c_int),
            "swm_index_host_free": ([C.POINTER(IndexHost)], None),
            "swm_index_write": ([C.POINTER(IndexHost), C.c_char_p, C.c_char_p], C.c_int),
            "swm_index_load": ([C.c_char_p, u32, C.POINTER(vp)], C.c_int),
            "swm_ctx_create": ([u32, vp, vp, vp, u64, vp, u64, C.POINTER(vp)], C.c_int),
            "swm_ctx_destroy": ([vp], None),
            "swm_map_batch": ([vp, vp, u64, vp, u64, u32, vp], C.c_int),
            "swm_map_batch_device": ([vp, vp, u64, vp, u64, u32, vp], C.c_int),
            "swm_map_batch_device_noalloc": ([vp, vp, u64, vp, u64, u32, vp], C.c_int),
            "swm_last_stats": ([vp, C.POINTER(Stats)], C.c_int),
            "swm_myers_device": ([vp, vp, vp, vp, vp, vp, u64, u32, vp, vp], C.c_int),
            "swm_synth_genome": ([vp, u64, u64, C.c_int, C.c_double], C.c_int),
            "swm_device_count": ([C.POINTER(C.c_int)], C.c_int),
            "swm_valu_probe": ([C.POINTER(C.c_double)], C.c_int),
            "swm_gather_probe": ([C.c_uint64, C.POINTER(C.c_double)], C.c_int),
            "swm_seed_slot": ([u32, u32, u32], u64),
            "swm_dev_alloc": ([C.POINTER(vp), u64], C.c_int),
            "swm_dev_free": ([vp], C.c_int),
            "swm_copy_h2d": ([vp, vp, u64], C.c_int),
            "swm_copy_d2h": ([vp, vp, u64], C.c_int),
            "swm_copy_d2d": ([vp, vp, u64], C.c_int),
            "swm_dev_memset": ([vp, C.c_int, u64], C.c_int),
            "swm_synth_reads": ([vp, u64, u64, u32, u32, C.c_double, C.c_double, C.c_double, u64,
                                 vp, u64, vp, vp, vp], C.c_int),
            "swm_ctx_set_option": ([vp, C.c_int, C.c_int64], C.c_int),
            "swm_ctx_get_option": ([vp, C.c_int, C.POINTER(C.c_int64)], C.c_int),
            "swm_autotune": ([vp, vp, u64, vp, u64, u32, u32, C.POINTER(C.c_int64)], C.c_int),
            "swm_server_create": ([vp, C.POINTER(ServerConfig), C.POINTER(vp)], C.c_int),
            "swm_server_submit": ([vp, vp, vp, u64, vp, C.POINTER(u64)], C.c_int),
            "swm_server_wait": ([vp, u64, C.POINTER(BatchInfo)], C.c_int),
            "swm_server_warmup": ([vp, vp, vp, u64], C.c_int),
            "swm_server_totals": ([vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)], C.c_int),
            "swm_server_destroy": ([vp], None),
            "swm_host_alloc": ([C.POINTER(vp), u64], C.c_int),
            "swm_host_free": ([vp], C.c_int),
            "swm_map_batch_records": ([vp, vp, u64, vp, u64, u32, vp, C.POINTER(Records)], C.c_int),
            "swm_records_free": ([C.POINTER(Records)], None),
            "swm_records_to_ref_blocks": ([C.POINTER(Records), u64, C.POINTER(RefBlocks)], C.c_int),
            "swm_ref_blocks_free": ([C.POINTER(RefBlocks)], None),
            "swm_sam_open": ([C.c_char_p, C.c_char_p, C.c_int, C.POINTER(vp)], C.c_int),
            "swm_sam_write_batch": ([vp, u64, C.POINTER(C.c_char_p), vp, vp, vp, u64, vp, C.POINTER(Records)],
                                    C.c_int),
            "swm_sam_close": ([vp], C.c_int),
            "swm_sam_set_threads": ([vp, C.c_int], C.c_int),
            "swm_seqfile_open": ([C.c_char_p], vp),
            "swm_seqfile_next": ([vp, C.POINTER(SeqRec)], C.c_int),
            "swm_seqfile_close": ([vp], None),
            "swm_seqrec_free": ([C.POINTER(SeqRec)], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib_handle = L
    return _lib_handle


def _check(fn, rc):
    if rc != SWM_OK:
        raise SwmError(fn, rc)


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


# ---------------------------------------------------------------- synthetic data

def synth_genome(n, seed=1, kind=GENOME_UNIFORM, frac=0.0):
    g = np.empty(n, dtype=np.uint8)
    _check("swm_synth_genome", lib().swm_synth_genome(_ptr(g), n, seed, kind, frac))
    return g


def synth_reads(genome, n_reads, len_min=150, len_max=150, err_min=0.025, err_max=0.025, indel_share=0.2,
                seed=7, stride=256):
    reads = np.zeros((n_reads, stride), dtype=np.uint8)
    lens = np.zeros(n_reads, dtype=np.uint32)
    tpos = np.zeros(n_reads, dtype=np.uint32)
    tstr = np.zeros(n_reads, dtype=np.uint8)
    _check("swm_synth_reads", lib().swm_synth_reads(_ptr(genome), genome.size, n_reads, len_min, len_max, err_min,
                                                    err_max, indel_share, seed, _ptr(reads), stride, _ptr(lens),
                                                    _ptr(tpos), _ptr(tstr)))
    return reads, lens, tpos, tstr


# ---------------------------------------------------------------- FASTA/FASTQ

def read_fastx(path):
    """Records of a FASTA/FASTQ file as the reference's kseq_read returns
    them: a list of (name, seq, qual) bytes, and the last return code
    (-1 at a clean end, -2 for a quality string of another length)."""
    L = lib()
    f = L.swm_seqfile_open(path.encode())
    if not f:
        raise SwmError("swm_seqfile_open", -2)
    rec = SeqRec()
    out = []
    try:
        while True:
            rc = L.swm_seqfile_next(f, C.byref(rec))
            if rc < 0:
                return out, rc
            out.append((C.string_at(rec.name, rec.name_len), C.string_at(rec.seq, rec.seq_len),
                        C.string_at(rec.qual, rec.qual_len)))
    finally:
        L.swm_seqfile_close(f)
        L.swm_seqrec_free(C.byref(rec))


# ---------------------------------------------------------------- index

class Index:
    """In-memory index: the exact content of the reference's index files."""

    def __init__(self, k, n_bases, base_2bits, base_3bits, hash_count, hash_offset_end, positions, pos_lseeds):
        self.k, self.n_bases = k, n_bases
        self.base_2bits, self.base_3bits = base_2bits, base_3bits
        self.hash_count, self.hash_offset_end = hash_count, hash_offset_end
        self.positions, self.pos_lseeds = positions, pos_lseeds
        self.build_ms = None

    def files(self):
        return {".base_2bits": self.base_2bits, ".base_3bits": self.base_3bits, ".hash_count": self.hash_count,
                ".hash_offset": self.hash_offset_end, ".positions": self.positions, ".pos_lseeds": self.pos_lseeds}

    def write(self, prefix, contig_name="chr1"):
        ih = IndexHost(self.k, self.n_bases, (self.base_2bits.size + 1) // 2,
                       self.base_2bits.ctypes.data_as(C.POINTER(C.c_uint32)),
                       self.base_3bits.ctypes.data_as(C.POINTER(C.c_uint32)),
                       self.hash_count.ctypes.data_as(C.POINTER(C.c_uint32)),
                       self.hash_offset_end.ctypes.data_as(C.POINTER(C.c_uint32)),
                       self.positions.ctypes.data_as(C.POINTER(C.c_uint32)),
                       self.pos_lseeds.ctypes.data_as(C.POINTER(C.c_uint16)), self.positions.size)
        _check("swm_index_write", lib().swm_index_write(C.byref(ih), prefix.encode(), contig_name.encode()))


def build_index(seq, k=12):
    """GPU index build of a concatenated sequence (uint8 array or bytes)."""
    s = np.frombuffer(seq, dtype=np.uint8) if isinstance(seq, (bytes, bytearray)) else np.ascontiguousarray(seq)
    ih = IndexHost()
    ms = C.c_double(0)
    _check("swm_index_build_mem", lib().swm_index_build_mem(_ptr(s), s.size, k, C.byref(ih), C.byref(ms)))
    try:
        nw, nb, m = ih.n_words, 1 << (2 * k), ih.n_pos

        def cp(p, n):
            return np.ctypeslib.as_array(p, (n,)).copy() if n else np.zeros(0, dtype=np.dtype(p._type_))
        ix = Index(k, ih.n_bases, cp(ih.base_2bits, 2 * nw), cp(ih.base_3bits, nw), cp(ih.hash_count, nb),
                   cp(ih.hash_offset_end, nb + 1), cp(ih.positions, m), cp(ih.pos_lseeds, m))
        ix.build_ms = ms.value
        return ix
    finally:
        lib().swm_index_host_free(C.byref(ih))


def index_build_files(fasta_path, prefix, k=12):
    _check("swm_index_build", lib().swm_index_build(fasta_path.encode(), prefix.encode(), k))


# ---------------------------------------------------------------- mapping

class Mapper:
    """Device-resident index context (one HIP stream)."""

    def __init__(self, handle, k):
        self._h = handle
        self.k = k

    @classmethod
    def from_index(cls, ix):
        h = C.c_void_p()
        _check("swm_ctx_create", lib().swm_ctx_create(ix.k, _ptr(ix.hash_count), _ptr(ix.positions),
                                                      _ptr(ix.pos_lseeds), ix.positions.size, _ptr(ix.base_2bits),
                                                      ix.base_2bits.size, C.byref(h)))
        return cls(h, ix.k)

    @classmethod
    def from_files(cls, prefix, k=12):
        h = C.c_void_p()
        _check("swm_index_load", lib().swm_index_load(prefix.encode(), k, C.byref(h)))
        return cls(h, k)

    def close(self):
        if self._h:
            lib().swm_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def map(self, reads, lens, e=5):
        """reads: (n, stride) uint8; lens: (n,) uint32.  Returns RESULT_DTYPE[n]."""
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        lens = np.ascontiguousarray(lens, dtype=np.uint32)
        n, stride = reads.shape
        out = np.zeros(n, dtype=RESULT_DTYPE)
        _check("swm_map_batch", lib().swm_map_batch(self._h, _ptr(reads), stride, _ptr(lens), n, e, _ptr(out)))
        return out

    def map_records(self, reads, lens, e=5, blocks=False):
        """swm_map_batch_records: (RESULT_DTYPE[n], RECORD_DTYPE[m], CIGAR text bytes), and with
        blocks=True also the put_cigar blocks' first-record indices (uint32)."""
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        lens = np.ascontiguousarray(lens, dtype=np.uint32)
        n, stride = reads.shape
        out = np.zeros(n, dtype=RESULT_DTYPE)
        r = Records()
        _check("swm_map_batch_records", lib().swm_map_batch_records(self._h, _ptr(reads), stride, _ptr(lens), n, e,
                                                                    _ptr(out), C.byref(r)))
        try:
            recs = np.frombuffer(C.string_at(r.rec, r.n * RECORD_DTYPE.itemsize), dtype=RECORD_DTYPE).copy() \
                if r.n else np.zeros(0, RECORD_DTYPE)
            text = C.string_at(r.cigar, r.text_bytes) if r.text_bytes else b""
            blk = np.frombuffer(C.string_at(r.block, r.n_blocks * 4), dtype=np.uint32).copy() \
                if r.n_blocks else np.zeros(0, np.uint32)
        finally:
            lib().swm_records_free(C.byref(r))
        return (out, recs, text, blk) if blocks else (out, recs, text)

    def map_device(self, d_reads, stride, d_lens, n, e, d_out):
        """Device pointers (ints), e.g. torch tensors' data_ptr()."""
        _check("swm_map_batch_device", lib().swm_map_batch_device(self._h, C.c_void_p(d_reads), stride,
                                                                  C.c_void_p(d_lens), n, e, C.c_void_p(d_out)))

    def map_device_noalloc(self, d_reads, stride, d_lens, n, e, d_out):
        """swm_map_batch_device_noalloc: the same mapping with no allocation
        (buffers carved from the library's fixed device scratch)."""
        _check("swm_map_batch_device_noalloc", lib().swm_map_batch_device_noalloc(
            self._h, C.c_void_p(d_reads), stride, C.c_void_p(d_lens), n, e, C.c_void_p(d_out)))

    def myers_device(self, d_read_words, d_lens, d_cread, d_wstart, d_rlen, n, e, d_err, d_loc):
        _check("swm_myers_device", lib().swm_myers_device(self._h, *(C.c_void_p(x) for x in (
            d_read_words, d_lens, d_cread, d_wstart, d_rlen)), n, e, C.c_void_p(d_err), C.c_void_p(d_loc)))

    def stats(self):
        s = Stats()
        _check("swm_last_stats", lib().swm_last_stats(self._h, C.byref(s)))
        return s.as_dict()

    def set_option(self, opt, value):
        _check("swm_ctx_set_option", lib().swm_ctx_set_option(self._h, opt, int(value)))

    def get_option(self, opt):
        v = C.c_int64(0)
        _check("swm_ctx_get_option", lib().swm_ctx_get_option(self._h, opt, C.byref(v)))
        return v.value

    def options(self):
        return {name: self.get_option(o) for o, name in OPTION_NAMES.items()}

    def autotune(self, reads, lens, e=5, reps=3):
        """Times the pipeline on a sample for each tunable's candidates and
        keeps the fastest; returns the chosen options."""
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        lens = np.ascontiguousarray(lens, dtype=np.uint32)
        n, stride = reads.shape
        ch = (C.c_int64 * N_OPTIONS)()
        _check("swm_autotune", lib().swm_autotune(self._h, _ptr(reads), stride, _ptr(lens), n, e, reps, ch))
        return {OPTION_NAMES[i + 1]: ch[i] for i in range(N_OPTIONS)}


def ref_blocks(recs, text, blocks, n_reads):
    """swm_records_to_ref_blocks: the reference's ref_pos_buffer (uint32[n_blocks, 258]),
    cigar_buffer (uint8[n_blocks, 4096]) and unmapped list (uint32) for one batch."""
    recs = np.ascontiguousarray(recs, dtype=RECORD_DTYPE)
    blocks = np.ascontiguousarray(blocks, dtype=np.uint32)
    tbuf = np.frombuffer(text + b"\0", dtype=np.uint8)
    r = Records(recs.size, recs.ctypes.data if recs.size else None, len(text), tbuf.ctypes.data,
                blocks.size, blocks.ctypes.data if blocks.size else None)
    b = RefBlocks()
    _check("swm_records_to_ref_blocks", lib().swm_records_to_ref_blocks(C.byref(r), n_reads, C.byref(b)))
    try:
        nb, nu = b.n_blocks, b.n_unmapped
        pos = np.frombuffer(C.string_at(b.ref_pos, nb * REF_POS_STRIDE * 4), np.uint32).reshape(nb, REF_POS_STRIDE) \
            if nb else np.zeros((0, REF_POS_STRIDE), np.uint32)
        cig = np.frombuffer(C.string_at(b.cigar, nb * REF_TEXT_STRIDE), np.uint8).reshape(nb, REF_TEXT_STRIDE) \
            if nb else np.zeros((0, REF_TEXT_STRIDE), np.uint8)
        unm = np.frombuffer(C.string_at(b.unmapped, nu * 4), np.uint32).copy() if nu else np.zeros(0, np.uint32)
    finally:
        lib().swm_ref_blocks_free(C.byref(b))
    return pos.copy(), cig.copy(), unm


def cigar_strings(recs, text):
    """The CIGAR text of each record."""
    return [text[int(o):int(o) + int(n)] for o, n in zip(recs["cigar_off"], recs["cigar_len"])]


class SamWriter:
    """swm_sam_*: the reference's SAM output (header, unmapped lines, record
    lines) for batches mapped with Mapper.map_records."""

    def __init__(self, path, index_prefix, gzip_level=1, threads=None):
        h = C.c_void_p()
        _check("swm_sam_open", lib().swm_sam_open(path.encode(), index_prefix.encode(), gzip_level, C.byref(h)))
        self._h = h
        if threads is not None:
            _check("swm_sam_set_threads", lib().swm_sam_set_threads(h, threads))

    def write(self, names, reads, quals, lens, recs, text):
        """names: list of bytes; reads / quals: (n, stride) uint8 (quals may be
        None); recs / text as returned by Mapper.map_records."""
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        lens = np.ascontiguousarray(lens, dtype=np.uint32)
        n, stride = reads.shape
        if quals is not None:
            quals = np.ascontiguousarray(quals, dtype=np.uint8)
            assert quals.shape == reads.shape
        recs = np.ascontiguousarray(recs, dtype=RECORD_DTYPE)
        tbuf = np.frombuffer(text + b"\0", dtype=np.uint8)
        arr = (C.c_char_p * max(n, 1))(*names)
        nlen = np.array([len(x) for x in names], dtype=np.uint32)
        r = Records(recs.size, recs.ctypes.data if recs.size else None, len(text), tbuf.ctypes.data, 0, None)
        _check("swm_sam_write_batch", lib().swm_sam_write_batch(self._h, n, arr, _ptr(nlen), _ptr(reads), _ptr(quals),
                                                                stride, _ptr(lens), C.byref(r)))

    def close(self):
        if self._h:
            h, self._h = self._h, None
            _check("swm_sam_close", lib().swm_sam_close(h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching serving loop over a Mapper's resident index: when
    one of n_slots streams is free it takes the oldest waiting batches that
    fit max_launch reads and the HBM budget and maps them in one launch
    (max_launch=0: one batch per launch)."""

    def __init__(self, mapper, e=5, stride=256, n_slots=4, hbm_budget=0, max_batch=65536, graph=True,
                 max_launch=262144):
        self.mapper = mapper
        cfg = ServerConfig(n_slots, e, stride, hbm_budget, max_batch, max_launch, 1 if graph else 0, 0)
        h = C.c_void_p()
        _check("swm_server_create", lib().swm_server_create(mapper._h, C.byref(cfg), C.byref(h)))
        self._h = h
        self._keep = {}

    def submit(self, reads, lens):
        """Non-blocking; returns a ticket.  The arrays are kept alive until wait()."""
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        lens = np.ascontiguousarray(lens, dtype=np.uint32)
        out = np.zeros(lens.size, dtype=RESULT_DTYPE)
        t = C.c_uint64(0)
        _check("swm_server_submit", lib().swm_server_submit(self._h, _ptr(reads), _ptr(lens), lens.size, _ptr(out),
                                                            C.byref(t)))
        self._keep[t.value] = (reads, lens, out)
        return t.value

    def warmup(self, reads, lens):
        """Learn capacities and capture graphs on every slot before traffic."""
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        lens = np.ascontiguousarray(lens, dtype=np.uint32)
        _check("swm_server_warmup", lib().swm_server_warmup(self._h, _ptr(reads), _ptr(lens), lens.size))

    def wait(self, ticket):
        """Blocks until the batch is done: (results, info dict)."""
        info = BatchInfo()
        rc = lib().swm_server_wait(self._h, ticket, C.byref(info))
        reads, lens, out = self._keep.pop(ticket)
        _check("swm_server_wait", rc)
        d = {f: getattr(info, f) for f, _ in BatchInfo._fields_ if f != "stats"}
        d["stats"] = info.stats.as_dict()
        return out, d

    def totals(self):
        v = [C.c_uint64(0) for _ in range(4)]
        _check("swm_server_totals", lib().swm_server_totals(self._h, *(C.byref(x) for x in v)))
        return dict(zip(("batches", "reads", "peak_bytes", "budget"), (x.value for x in v)))

    def close(self):
        if self._h:
            lib().swm_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def valu_peak():
    """Measured int32 VALU lane-ops/s of the device."""
    v = C.c_double(0)
    _check("swm_valu_probe", lib().swm_valu_probe(C.byref(v)))
    return v.value


def gather_peak(array_bytes=512 << 20):
    """Measured random 8-byte gathers per second (one line read each) over a
    device array of array_bytes."""
    v = C.c_double(0)
    _check("swm_gather_probe", lib().swm_gather_probe(int(array_bytes), C.byref(v)))
    return v.value


def device_count():
    """HIP devices visible to the library (0 on a CPU-only host)."""
    n = C.c_int(0)
    lib().swm_device_count(C.byref(n))
    return n.value


class DeviceBuffer:
    """HBM buffer owned through the library's HIP runtime (no torch in the
    process: torch ships its own HIP runtime build)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check("swm_dev_alloc", lib().swm_dev_alloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _check("swm_copy_h2d", lib().swm_copy_h2d(C.c_void_p(b.ptr), C.c_void_p(a.ctypes.data), a.nbytes))
        return b

    def zero(self):
        _check("swm_dev_memset", lib().swm_dev_memset(C.c_void_p(self.ptr), 0, self.nbytes))
        return self

    def to_array(self, dtype, shape=None):
        dt = np.dtype(dtype)
        out = np.empty(self.nbytes // dt.itemsize, dtype=dt)
        _check("swm_copy_d2h", lib().swm_copy_d2h(C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr), out.nbytes))
        return out.reshape(shape) if shape is not None else out

    def free(self):
        if self.ptr:
            lib().swm_dev_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """Page-locked host memory as a numpy array (``.a``): batches handed to
    the serving loop from pinned memory move over PCIe by DMA."""

    def __init__(self, shape, dtype):
        dt = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dt.itemsize
        p = C.c_void_p()
        _check("swm_host_alloc", lib().swm_host_alloc(C.byref(p), max(nbytes, 1)))
        self.ptr = p.value
        buf = (C.c_uint8 * max(nbytes, 1)).from_address(self.ptr)
        self.a = np.frombuffer(buf, dtype=np.uint8, count=nbytes).view(dt).reshape(shape)

    def free(self):
        if self.ptr:
            self.a = None
            lib().swm_host_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

"""ctypes bindings of the oracle (oracle/_build/libswm_oracle.so).

Test infrastructure only: the oracle is the checker, never the thing measured.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libswm_oracle.so")

HIT_DTYPE = np.dtype([("pos", "<u4"), ("strand", "u1"), ("dist", "u1"), ("pad", "<u2"), ("count", "<u4")])
REC_DTYPE = np.dtype([("read", "<u4"), ("strand", "u1"), ("dist", "u1"), ("block", "<u2"), ("loc", "<u4"),
                      ("text_len", "<u4"), ("primary", "u1"), ("pad", "u1", (3,)), ("cigar", "S160")])
CAND_DTYPE = np.dtype([("read", "<u4"), ("strand", "u1"), ("pad0", "u1"), ("batch", "<u2"), ("idx", "<u4"),
                       ("cand", "<u4"), ("win_start", "<u4"), ("ref_len", "<u4"), ("err", "<i4"), ("loc", "<i4")])


class OrcIndex(C.Structure):
    _fields_ = [("k", C.c_uint32), ("n_bases", C.c_uint64), ("n_words", C.c_uint64),
                ("base_2bits", C.POINTER(C.c_uint32)), ("base_3bits", C.POINTER(C.c_uint32)),
                ("hash_count", C.POINTER(C.c_uint32)), ("hash_offset_end", C.POINTER(C.c_uint32)),
                ("positions", C.POINTER(C.c_uint32)), ("pos_lseeds", C.POINTER(C.c_uint16)),
                ("n_pos", C.c_uint64)]


class OrcRef(C.Structure):
    _fields_ = [("hash_offset", C.c_void_p), ("positions", C.c_void_p), ("pos_lseeds", C.c_void_p),
                ("base_2bits", C.c_void_p), ("n_2bits_words", C.c_uint64), ("ref_length", C.c_uint64),
                ("n_pos", C.c_uint64), ("k", C.c_uint32)]


_LIBS = {}


def mutant_so(m):
    """The oracle built with planted mistake m (ORC_MUTANT in swm_oracle.c)."""
    return os.path.join(ROOT, "oracle", "_build", f"libswm_oracle_m{m}.so")


def lib(path=ORACLE_SO):
    if path not in _LIBS:
        if not os.path.exists(path):
            import subprocess
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, stdout=subprocess.DEVNULL)
        L = C.CDLL(path)
        vp, u32, u64 = C.c_void_p, C.c_uint32, C.c_uint64
        L.orc_index_build.argtypes = [vp, u64, u32, C.POINTER(OrcIndex)]
        L.orc_index_build.restype = C.c_int
        L.orc_index_free.argtypes = [C.POINTER(OrcIndex)]
        L.orc_map.argtypes = [C.POINTER(OrcRef), vp, u64, vp, u64, u32, vp, C.POINTER(u64), C.c_int]
        L.orc_map.restype = C.c_int
        L.orc_map_trace.argtypes = [C.POINTER(OrcRef), vp, u64, vp, u64, u32, vp, C.POINTER(vp), C.POINTER(u64),
                                    C.POINTER(vp), C.POINTER(u64)]
        L.orc_map_trace.restype = C.c_int
        L.orc_free.argtypes = [vp]
        L.orc_encode_read.argtypes = [vp, u32, u32, vp, vp]
        L.orc_encode_read.restype = C.c_int
        L.orc_reverse_read.argtypes = [vp, vp, u32]
        L.orc_myers.argtypes = [vp, vp, vp, u32, u32, u32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.orc_myers64.argtypes = L.orc_myers.argtypes
        _LIBS[path] = L
    return _LIBS[path]


def _p(a):
    return C.c_void_p(a.ctypes.data)


class OracleIndex:
    """Index built by the oracle's restatement of serial_index.c."""

    def __init__(self, seq, k=12):
        s = np.frombuffer(seq, dtype=np.uint8) if isinstance(seq, (bytes, bytearray)) else np.ascontiguousarray(seq)
        ix = OrcIndex()
        rc = lib().orc_index_build(_p(s), s.size, k, C.byref(ix))
        if rc != 0:
            raise ValueError(f"orc_index_build rc={rc}")
        try:
            nw, nb, m = ix.n_words, 1 << (2 * k), ix.n_pos
            self.k, self.n_bases = k, ix.n_bases
            self.base_2bits = np.ctypeslib.as_array(ix.base_2bits, (2 * nw,)).copy()
            self.base_3bits = np.ctypeslib.as_array(ix.base_3bits, (nw,)).copy()
            self.hash_count = np.ctypeslib.as_array(ix.hash_count, (nb,)).copy()
            self.hash_offset_end = np.ctypeslib.as_array(ix.hash_offset_end, (nb + 1,)).copy()
            self.positions = np.ctypeslib.as_array(ix.positions, (m,)).copy() if m else np.zeros(0, np.uint32)
            self.pos_lseeds = np.ctypeslib.as_array(ix.pos_lseeds, (m,)).copy() if m else np.zeros(0, np.uint16)
        finally:
            lib().orc_index_free(C.byref(ix))

    def files(self):
        return {".base_2bits": self.base_2bits, ".base_3bits": self.base_3bits, ".hash_count": self.hash_count,
                ".hash_offset": self.hash_offset_end, ".positions": self.positions, ".pos_lseeds": self.pos_lseeds}


class OracleRef:
    """The mapper's in-memory view of an index (load_ref_info, mapper_mpe.c:198-314)."""

    def __init__(self, k, hash_count, positions, pos_lseeds, base_2bits, so=ORACLE_SO):
        self.k = k
        self.L = lib(so)
        nb = 1 << (2 * k)
        off = np.zeros(nb + 1, dtype=np.uint64)
        np.cumsum(hash_count.astype(np.uint64), out=off[1:])
        self.hash_offset = (off & 0xFFFFFFFF).astype(np.uint32)
        self.positions = np.ascontiguousarray(positions, dtype=np.uint32)
        self.pos_lseeds = np.ascontiguousarray(pos_lseeds, dtype=np.uint16)
        nw = base_2bits.size
        self.base_2bits = np.zeros(nw + 128, dtype=np.uint32)
        self.base_2bits[:nw] = base_2bits
        self.ref = OrcRef(self.hash_offset.ctypes.data, self.positions.ctypes.data, self.pos_lseeds.ctypes.data,
                          self.base_2bits.ctypes.data, nw, 16 * nw * 4, self.positions.size, k)

    @classmethod
    def from_index(cls, ix, so=ORACLE_SO):
        return cls(ix.k, ix.hash_count, ix.positions, ix.pos_lseeds, ix.base_2bits, so=so)

    def map(self, reads, lens, e, threads=0):
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        lens = np.ascontiguousarray(lens, dtype=np.uint32)
        n, stride = reads.shape
        hits = np.zeros(n, dtype=HIT_DTYPE)
        nv = C.c_uint64(0)
        rc = self.L.orc_map(C.byref(self.ref), _p(reads), stride, _p(lens), n, e, _p(hits), C.byref(nv), threads)
        if rc != 0:
            raise ValueError(f"orc_map rc={rc}")
        return hits, nv.value

    def trace(self, reads, lens, e):
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        lens = np.ascontiguousarray(lens, dtype=np.uint32)
        n, stride = reads.shape
        hits = np.zeros(n, dtype=HIT_DTYPE)
        rp, nr, cp, ncd = C.c_void_p(), C.c_uint64(), C.c_void_p(), C.c_uint64()
        rc = self.L.orc_map_trace(C.byref(self.ref), _p(reads), stride, _p(lens), n, e, _p(hits), C.byref(rp),
                                 C.byref(nr), C.byref(cp), C.byref(ncd))
        if rc != 0:
            raise ValueError(f"orc_map_trace rc={rc}")
        try:
            recs = np.frombuffer(C.string_at(rp, nr.value * REC_DTYPE.itemsize), dtype=REC_DTYPE).copy() \
                if nr.value else np.zeros(0, REC_DTYPE)
            cands = np.frombuffer(C.string_at(cp, ncd.value * CAND_DTYPE.itemsize), dtype=CAND_DTYPE).copy() \
                if ncd.value else np.zeros(0, CAND_DTYPE)
        finally:
            self.L.orc_free(rp)
            self.L.orc_free(cp)
        return hits, recs, cands


def encode_read(read, k=12):
    r = np.frombuffer(read, dtype=np.uint8) if isinstance(read, (bytes, bytearray)) else np.ascontiguousarray(read)
    buf = np.zeros(r.size + 64, dtype=np.uint8)
    buf[:r.size] = r
    words = np.zeros(32, dtype=np.uint32)
    seeds = np.zeros(128, dtype=np.uint32)
    n = lib().orc_encode_read(_p(buf), r.size, k, _p(words), _p(seeds))
    return words, seeds[:n]


def reverse_read(read):
    r = np.frombuffer(read, dtype=np.uint8) if isinstance(read, (bytes, bytearray)) else np.ascontiguousarray(read)
    out = np.zeros(r.size, dtype=np.uint8)
    lib().orc_reverse_read(_p(out), _p(r), r.size)
    return out


def myers(read_words, win_lo, win_hi, L, ref_len, e, width=32):
    err, loc = C.c_int32(), C.c_int32()
    f = lib().orc_myers if width == 32 else lib().orc_myers64
    f(_p(np.ascontiguousarray(read_words, np.uint32)), _p(np.ascontiguousarray(win_lo, np.uint32)),
      _p(np.ascontiguousarray(win_hi, np.uint32)), L, ref_len, e, C.byref(err), C.byref(loc))
    return err.value, loc.value


def block_starts(orecs):
    """Indices of the oracle records that open a put_cigar block: a new
    (read, strand) or a new flush block number within it."""
    if orecs.size == 0:
        return np.zeros(0, np.uint32)
    key = (orecs["read"].astype(np.int64) * 2 + orecs["strand"]) * 65536 + orecs["block"]
    return np.flatnonzero(np.r_[True, key[1:] != key[:-1]]).astype(np.uint32)

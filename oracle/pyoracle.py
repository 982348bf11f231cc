"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (libbpx.so and the
bpx package) never imports it.  See oracle.h for what each function restates
(reference file:line) and tests/ for what pins it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("BPX_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built (make -C oracle)")
_lib = C.CDLL(LIB_PATH)

HSP_DTYPE = np.dtype([("hsp_q", "<i4"), ("hsp_s", "<i4"), ("query_offset", "<i4"),
                      ("subject_offset", "<i4"), ("hsp_len", "<i4"), ("score", "<i4")])


class Params(C.Structure):
    _fields_ = [("word_length", C.c_int32), ("threshold", C.c_int32), ("window", C.c_int32),
                ("x_dropoff", C.c_int32), ("cutoff", C.c_int32), ("diag_mask", C.c_int32)]


class Lookup(C.Structure):
    _fields_ = [("word_length", C.c_int32), ("nwords", C.c_uint32),
                ("cell_start", C.POINTER(C.c_uint32)), ("entries", C.POINTER(C.c_uint16)),
                ("max_cell", C.c_uint32)]


class Diag(C.Structure):
    _fields_ = [("last_hit", C.c_int32), ("flag", C.c_int32)]


class ExtLeft(C.Structure):
    _fields_ = [("maxscore", C.c_int32), ("length", C.c_int32)]


class ExtRight(C.Structure):
    _fields_ = [("maxscore", C.c_int32), ("length", C.c_int32), ("s_last_off", C.c_int32)]


class ExtTwo(C.Structure):
    _fields_ = [("hsp_q", C.c_int32), ("hsp_s", C.c_int32), ("hsp_len", C.c_int32),
                ("score", C.c_int32), ("right_extend", C.c_int32), ("s_last_off", C.c_int32)]


_vp = C.c_void_p


def _sig(name, res, *args):
    fn = getattr(_lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


_sig("or_lookup_build", C.c_int, _vp, C.c_int32, _vp, C.c_int32, C.c_int32, C.POINTER(Lookup))
_sig("or_lookup_free", None, C.POINTER(Lookup))
_sig("or_diag_length", C.c_int32, C.c_int32, C.c_int32)
_sig("or_extend_left", ExtLeft, _vp, _vp, _vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32)
_sig("or_extend_right", ExtRight, _vp, _vp, _vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
     C.c_int32, C.c_int32)
_sig("or_extend_two_hit", ExtTwo, _vp, _vp, C.c_int32, _vp, C.c_int32, C.c_int32, C.c_int32,
     C.c_int32, C.c_int32, C.c_int32)
_sig("or_scan_subject", C.c_int32, C.POINTER(Lookup), _vp, C.c_int32, _vp, C.POINTER(Params), _vp,
     C.c_int32, C.POINTER(Diag), C.POINTER(C.c_int32), _vp, C.c_int32, _vp)
_sig("or_run_db", C.c_int, C.POINTER(Lookup), _vp, C.c_int32, _vp, C.POINTER(Params), _vp, _vp, _vp,
     C.c_uint32, C.c_int, _vp, _vp, C.POINTER(_vp))
_sig("or_free", None, _vp)
_sig("or_blosum62", C.POINTER(C.c_int8))


def _p(a):
    return a.ctypes.data_as(_vp)


def blosum62() -> np.ndarray:
    return np.ctypeslib.as_array(_lib.or_blosum62(), shape=(28 * 28,)).copy().reshape(28, 28)


_M = None


def _matrix(m):
    global _M
    if m is None:
        if _M is None:
            _M = np.ascontiguousarray(blosum62().reshape(-1), dtype=np.int8)
        return _M
    return np.ascontiguousarray(m, dtype=np.int8).reshape(-1)


def diag_length(qlen: int, window: int = 40) -> int:
    return int(_lib.or_diag_length(qlen, window))


def params(qlen: int, word_length=3, threshold=11, window=40, x_dropoff=16, cutoff=41, diag_mask=0):
    if diag_mask == 0:
        diag_mask = diag_length(qlen, window) - 1
    return Params(word_length, threshold, window, x_dropoff, cutoff, diag_mask)


class OracleLookup:
    """Brute-force neighbourhood table (owned; freed on close)."""

    def __init__(self, query, threshold=11, word_length=3, matrix=None):
        self.query = np.ascontiguousarray(query, dtype=np.uint8)
        self.matrix = _matrix(matrix)
        self.lk = Lookup()
        rc = _lib.or_lookup_build(_p(self.query), len(self.query), _p(self.matrix), threshold,
                                  word_length, C.byref(self.lk))
        if rc != 0:
            raise ValueError("or_lookup_build failed")

    def csr(self):
        n = self.lk.nwords
        cs = np.ctypeslib.as_array(self.lk.cell_start, shape=(n + 1,)).copy()
        ne = int(cs[-1])
        ent = (np.ctypeslib.as_array(self.lk.entries, shape=(ne,)).copy() if ne
               else np.zeros(0, dtype=np.uint16))
        return cs, ent

    def close(self):
        if self.lk.cell_start:
            _lib.or_lookup_free(C.byref(self.lk))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def extend_left(subject, query, s_off, q_off, dropoff, maxscore=0, matrix=None):
    s, q, m = np.ascontiguousarray(subject, np.uint8), np.ascontiguousarray(query, np.uint8), _matrix(matrix)
    r = _lib.or_extend_left(_p(m), _p(s), _p(q), s_off, q_off, dropoff, maxscore)
    return r.maxscore, r.length


def extend_right(subject, query, s_off, q_off, dropoff, maxscore=0, matrix=None):
    s, q, m = np.ascontiguousarray(subject, np.uint8), np.ascontiguousarray(query, np.uint8), _matrix(matrix)
    r = _lib.or_extend_right(_p(m), _p(s), _p(q), s_off, q_off, dropoff, maxscore, len(s), len(q))
    return r.maxscore, r.length, r.s_last_off


def extend_two_hit(subject, query, s_left_off, s_right_off, q_right_off, dropoff, word=3, matrix=None):
    s, q, m = np.ascontiguousarray(subject, np.uint8), np.ascontiguousarray(query, np.uint8), _matrix(matrix)
    r = _lib.or_extend_two_hit(_p(m), _p(s), len(s), _p(q), len(q), s_left_off, s_right_off, q_right_off,
                               dropoff, word)
    return {"hsp_q": r.hsp_q, "hsp_s": r.hsp_s, "hsp_len": r.hsp_len, "score": r.score,
            "right_extend": r.right_extend, "s_last_off": r.s_last_off}


def run_db(query, residues, offsets, lengths, prm: Params | None = None, nthreads=1, matrix=None,
           lookup: OracleLookup | None = None):
    """Scans every subject; returns (counters[n,3], rec_offsets[n+1], records)."""
    q = np.ascontiguousarray(query, np.uint8)
    if prm is None:
        prm = params(len(q))
    m = _matrix(matrix)
    own = lookup is None
    lk = lookup if lookup is not None else OracleLookup(q, prm.threshold, prm.word_length, m)
    try:
        res = np.ascontiguousarray(residues, np.uint8)
        off = np.ascontiguousarray(offsets, np.uint64)
        ln = np.ascontiguousarray(lengths, np.uint32)
        n = len(ln)
        counters = np.zeros((max(n, 1), 3), dtype=np.int32)
        recoff = np.zeros(n + 1, dtype=np.uint64)
        recp = _vp()
        rc = _lib.or_run_db(C.byref(lk.lk), _p(q), len(q), _p(m), C.byref(prm), _p(res), _p(off), _p(ln), n,
                            nthreads, _p(counters), _p(recoff), C.byref(recp))
        if rc != 0:
            raise RuntimeError("or_run_db failed")
        total = int(recoff[-1])
        if total:
            recs = np.ctypeslib.as_array(C.cast(recp, C.POINTER(C.c_int32)), shape=(total * 6,)).copy()
            recs = recs.view(HSP_DTYPE)
        else:
            recs = np.zeros(0, dtype=HSP_DTYPE)
        _lib.or_free(recp)
        return counters[:n], recoff, recs
    finally:
        if own:
            lk.close()


def scan_subject_persistent(query, subjects, prm: Params | None = None, matrix=None):
    """Scans subjects one after another with ONE diagonal table and the NCBI
    running offset (never reset between subjects); returns per-subject
    (counters, records) lists."""
    q = np.ascontiguousarray(query, np.uint8)
    if prm is None:
        prm = params(len(q))
    m = _matrix(matrix)
    lk = OracleLookup(q, prm.threshold, prm.word_length, m)
    dlen = prm.diag_mask + 1
    diag = (Diag * dlen)()
    for i in range(dlen):
        diag[i].last_hit = -prm.window
        diag[i].flag = 0
    off = C.c_int32(prm.window)
    out = []
    try:
        for s in subjects:
            s = np.ascontiguousarray(s, np.uint8)
            cap = 4096
            buf = np.zeros(cap, dtype=HSP_DTYPE)
            cnt = np.zeros(3, dtype=np.int32)
            n = _lib.or_scan_subject(C.byref(lk.lk), _p(q), len(q), _p(m), C.byref(prm), _p(s), len(s), diag,
                                     C.byref(off), _p(buf), cap, _p(cnt))
            if n > cap:
                raise RuntimeError("record buffer too small")
            out.append((cnt.copy(), buf[:n].copy()))
        return out, off.value
    finally:
        lk.close()

"""Python binding of libbpx.so, the gfx950 blastp word finder (include/bpx.h).

The host-side mirror of the reference's GPU boundary
(ncbi_blast_files/gpu_cpu_common.h:40-98): a database volume is loaded once
into a resident device arena (bpx_db_load), a query's lookup table is built
and staged once (bpx_query_prepare), and a scan runs the word finder over a
range of subjects and returns the GPUBlastInitHitList-shaped records plus the
per-subject d_Hits counters (gpu_blastp_kernel.cu:264-276, 310-312).

Everything here is plumbing over the C ABI; the compute is in the HIP
kernels.  Importing this module never falls back to a CPU path: if the
shared library is missing the import raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libbpx.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run __graft_entry__.build() "
        "(make -C blastp-cuda_amd) first")
_lib = C.CDLL(LIB_PATH)

ALPHABET = 28
STDAA = "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ"

OUT_SUBJECT_STATS = 1
OUT_GRAPH = 2
OUT_PHASES = 4

HSP_DTYPE = np.dtype([("hsp_q", "<i4"), ("hsp_s", "<i4"), ("query_offset", "<i4"),
                      ("subject_offset", "<i4"), ("hsp_len", "<i4"), ("score", "<i4")])
STATS_DTYPE = np.dtype([("hits", "<i4"), ("successful_extensions", "<i4"),
                        ("hits_extended", "<i4")])


class BpxError(RuntimeError):
    pass


class WordParams(C.Structure):
    _fields_ = [("word_length", C.c_int32), ("threshold", C.c_int32), ("window", C.c_int32),
                ("x_dropoff", C.c_int32), ("cutoff_score", C.c_int32), ("diag_mask", C.c_int32)]


class Launch(C.Structure):
    _fields_ = [("waves_per_group", C.c_int32), ("groups_per_cu", C.c_int32),
                ("subjects_per_item", C.c_int32), ("queue_capacity", C.c_int32),
                ("subject_buffer", C.c_int32), ("claims_per_share", C.c_int32)]


class BatchPolicy(C.Structure):
    _fields_ = [("group_residues", C.c_int32), ("group_queries", C.c_int32), ("group_diag_slots", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("max_batch", C.c_int32), ("max_inflight", C.c_int32), ("max_wait_us", C.c_int32),
                ("diag_budget", C.c_int32), ("hbm_budget", C.c_int64), ("params", WordParams),
                ("policy", BatchPolicy), ("launch", Launch)]


class TuneTrial(C.Structure):
    _fields_ = [("policy", BatchPolicy), ("launch", Launch), ("sample_ms", C.c_float)]


class TuneResult(C.Structure):
    _fields_ = [("policy", BatchPolicy), ("launch", Launch), ("sample_ms", C.c_float), ("candidates", C.c_int32)]


class PhaseStats(C.Structure):
    _fields_ = [("waves", C.c_uint32), ("classes", C.c_uint32), ("span_ms", C.c_double),
                ("wave_ms_mean", C.c_double), ("wave_ms_min", C.c_double), ("stage_ms_mean", C.c_double),
                ("group_ms_mean", C.c_double), ("subjects_staged", C.c_uint64), ("bytes_staged", C.c_uint64),
                ("flush_ms_mean", C.c_double), ("extensions", C.c_uint64), ("busy_fraction", C.c_double)]


class UngappedStats(C.Structure):
    _fields_ = [("lookup_hits", C.c_int64), ("num_seqs_lookup_hits", C.c_int64),
                ("init_extends", C.c_int64), ("good_init_extends", C.c_int64),
                ("num_seqs_passed", C.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class RequestInfo(C.Structure):
    _fields_ = [("num_records", C.c_uint64), ("stats", UngappedStats), ("latency_ms", C.c_double),
                ("queue_ms", C.c_double), ("batch_size", C.c_uint32), ("reserved", C.c_uint32)]


class SynthSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("num_subjects", C.c_uint32),
                ("len_median", C.c_double), ("len_sigma", C.c_double),
                ("len_min", C.c_uint32), ("len_max", C.c_uint32),
                ("homolog_count", C.c_uint32), ("homolog_fraction", C.c_double),
                ("identity_min", C.c_double), ("identity_max", C.c_double),
                ("indel_min", C.c_double), ("indel_max", C.c_double),
                ("biased_fraction", C.c_double), ("dirichlet_alpha", C.c_double),
                ("motif_min", C.c_uint32), ("motif_max", C.c_uint32)]


_vp = C.c_void_p
_u8p = C.POINTER(C.c_uint8)


def _sig(name, res, *args):
    fn = getattr(_lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


_sig("bpx_available", C.c_int)
_sig("bpx_last_error", C.c_char_p)
_sig("bpx_context_create", C.c_int, C.c_int, C.POINTER(_vp))
_sig("bpx_context_destroy", C.c_int, _vp)
_sig("bpx_context_info", C.c_int, _vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64))
_sig("bpx_db_load", C.c_int, _vp, _vp, _vp, _vp, C.c_uint32, C.POINTER(_vp))
_sig("bpx_db_info", C.c_int, _vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64))
_sig("bpx_db_destroy", C.c_int, _vp)
_sig("bpx_query_prepare", C.c_int, _vp, _vp, C.c_int32, _vp, C.POINTER(WordParams), C.POINTER(_vp))
_sig("bpx_query_info", C.c_int, _vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
     C.POINTER(C.c_int64))
_sig("bpx_query_destroy", C.c_int, _vp)
_sig("bpx_context_limit", C.c_int, _vp, C.c_int64)
_sig("bpx_context_trim", C.c_int, _vp)
_sig("bpx_query_footprint", C.c_int, _vp, _vp, C.c_int32, C.POINTER(WordParams), C.POINTER(Launch),
     C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(C.c_size_t))
_sig("bpx_check_memory", C.c_int, _vp, _vp, _vp, C.c_uint32, C.c_uint32, C.POINTER(Launch),
     C.POINTER(C.c_size_t), C.POINTER(C.c_size_t))
_sig("bpx_scan", C.c_int, _vp, _vp, _vp, C.c_uint32, C.c_uint32, C.POINTER(Launch), C.c_uint32,
     C.POINTER(_vp))
_sig("bpx_scan_noalloc", C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_uint64, _vp, C.POINTER(C.c_uint64),
     C.POINTER(UngappedStats))
_sig("bpx_result_wait", C.c_int, _vp)
_sig("bpx_result_timing", C.c_int, _vp, C.POINTER(C.c_float), C.POINTER(C.c_float))
_sig("bpx_result_summary", C.c_int, _vp, C.POINTER(C.c_uint64), C.POINTER(UngappedStats))
_sig("bpx_result_records", C.c_int, _vp, _vp, _vp, _vp, C.c_uint64)
_sig("bpx_result_subject_stats", C.c_int, _vp, _vp, C.c_uint64)
_sig("bpx_result_query_stats", C.c_int, _vp, C.POINTER(UngappedStats), C.c_uint32)
_sig("bpx_batch_prepare", C.c_int, _vp, C.c_uint32, _vp, _vp, _vp, C.POINTER(WordParams), _vp, _vp,
     C.POINTER(BatchPolicy), C.POINTER(_vp))
_sig("bpx_batch_info", C.c_int, _vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_int64))
_sig("bpx_batch_destroy", C.c_int, _vp)
_sig("bpx_scan_batch", C.c_int, _vp, _vp, _vp, C.c_uint32, C.c_uint32, C.POINTER(Launch), C.c_uint32,
     C.POINTER(_vp))
_sig("bpx_result_destroy", C.c_int, _vp)
_sig("bpx_result_replay", C.c_int, _vp)
_sig("bpx_result_phases", C.c_int, _vp, C.POINTER(PhaseStats))
_sig("bpx_result_ready", C.c_int, _vp)
_sig("bpx_server_create", C.c_int, _vp, _vp, C.POINTER(ServerConfig), C.POINTER(_vp))
_sig("bpx_server_submit", C.c_int, _vp, _vp, C.c_int32, C.POINTER(C.c_uint64))
_sig("bpx_server_wait", C.c_int, _vp, C.c_uint64, C.c_int32, C.POINTER(RequestInfo))
_sig("bpx_server_records", C.c_int, _vp, C.c_uint64, _vp, _vp, C.c_uint64)
_sig("bpx_server_counters", C.c_int, _vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64))
_sig("bpx_server_destroy", C.c_int, _vp)
_sig("bpx_volume_read", C.c_int, C.c_char_p, C.POINTER(_vp))
_sig("bpx_volume_info", C.c_int, _vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
     C.POINTER(C.c_char_p))
_sig("bpx_volume_arrays", C.c_int, _vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp))
_sig("bpx_volume_destroy", C.c_int, _vp)
_sig("bpx_volume_write", C.c_int, C.c_char_p, C.c_char_p, _vp, _vp, _vp, C.c_uint32)
_sig("bpx_db_volumes", C.c_int, C.c_char_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(C.c_uint32))
_sig("bpx_db_load_volumes", C.c_int, _vp, C.c_char_p, C.POINTER(_vp))
_sig("bpx_volume_last_error", C.c_char_p)
_sig("bpx_gpu_info", C.c_int, C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32))
_sig("bpx_gpu_volume_read", C.c_int, C.c_char_p, C.c_uint32, C.POINTER(_vp))
_sig("bpx_db_features", C.c_int, _vp, _vp, _vp, C.c_int32, _vp)
_sig("bpx_db_set_cost", C.c_int, _vp, _vp)
_sig("bpx_costmodel_create", C.c_int, C.c_int, C.c_uint64, C.POINTER(_vp))
_sig("bpx_costmodel_params", C.c_int, _vp, _vp, C.c_int)
_sig("bpx_costmodel_train", C.c_int, _vp, _vp, _vp, C.c_uint32, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _vp)
_sig("bpx_costmodel_predict", C.c_int, _vp, _vp, C.c_uint32, _vp)
_sig("bpx_costmodel_grad", C.c_int, _vp, _vp, _vp, C.c_uint32, C.POINTER(C.c_float), _vp)
_sig("bpx_costmodel_last_error", C.c_char_p)
_sig("bpx_costmodel_destroy", C.c_int, _vp)
_sig("bpx_autotune_batch", C.c_int, _vp, C.c_uint32, _vp, _vp, _vp, C.POINTER(WordParams), _vp, _vp, _vp,
     C.c_uint32, C.c_uint32, C.POINTER(TuneResult), C.POINTER(TuneTrial), C.c_uint32, C.POINTER(C.c_uint32))
_sig("bpx_blosum62", C.POINTER(C.c_int8))
_sig("bpx_table_build", C.c_int, _vp, C.c_int32, _vp, C.c_int32, C.c_int32, _vp, _vp, C.c_uint64,
     C.POINTER(C.c_uint64))
_sig("bpx_synth_query", C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, _vp)
_sig("bpx_synth_query_lengths", C.c_int, C.c_uint64, C.c_uint32, C.c_int, C.c_double, C.c_double,
     C.c_double, C.c_double, _vp)
_sig("bpx_synth_db_generate", C.c_int, C.POINTER(SynthSpec), _vp, _vp, C.c_uint32, C.c_int,
     C.POINTER(_vp))
_sig("bpx_synth_db_arrays", C.c_int, _vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp),
     C.POINTER(C.c_uint32), C.POINTER(C.c_uint64))
_sig("bpx_synth_db_kinds", C.c_int, _vp, C.POINTER(_vp))
_sig("bpx_synth_db_sha256", C.c_int, _vp, C.c_char_p)
_sig("bpx_synth_db_free", C.c_int, _vp)
_sig("bpx_sha256_hex", C.c_int, _vp, C.c_uint64, C.c_char_p)


def _check(rc):
    if rc != 0:
        raise BpxError(f"bpx error {rc}: {_lib.bpx_last_error().decode(errors='replace')}")


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def available() -> bool:
    return bool(_lib.bpx_available())


def blosum62() -> np.ndarray:
    p = _lib.bpx_blosum62()
    return np.ctypeslib.as_array(p, shape=(ALPHABET * ALPHABET,)).copy().reshape(ALPHABET, ALPHABET)


def encode(seq: str) -> np.ndarray:
    """Letters -> ncbistdaa codes."""
    table = {c: i for i, c in enumerate(STDAA)}
    return np.array([table[c] for c in seq.upper()], dtype=np.uint8)


def default_params(qlen: int | None = None, **kw) -> WordParams:
    """blastp defaults: word 3, T 11, window 40, X-drop 16, cutoff 41 (raw);
    diag_mask 0 lets the library derive next_pow2(qlen + window) - 1."""
    p = WordParams(word_length=3, threshold=11, window=40, x_dropoff=16, cutoff_score=41, diag_mask=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def table_build(query: np.ndarray, threshold=11, word_length=3, matrix=None):
    """Host-side lookup table as (cell_start, entries) CSR over all words."""
    q = np.ascontiguousarray(query, dtype=np.uint8)
    m = None if matrix is None else np.ascontiguousarray(matrix, dtype=np.int8)
    nwords = 1 << (5 * word_length)
    cs = np.zeros(nwords + 1, dtype=np.uint32)
    n = C.c_uint64(0)
    rc = _lib.bpx_table_build(_ptr(q), len(q), _ptr(m), threshold, word_length, _ptr(cs), None, 0,
                              C.byref(n))
    ent = np.zeros(max(int(n.value), 1), dtype=np.uint16)
    _check(_lib.bpx_table_build(_ptr(q), len(q), _ptr(m), threshold, word_length, _ptr(cs), _ptr(ent),
                                len(ent), C.byref(n)))
    del rc
    return cs, ent[: int(n.value)]


# ---- synthetic workloads ----------------------------------------------------

def synth_query(seed: int, index: int, length: int) -> np.ndarray:
    out = np.zeros(max(length, 1), dtype=np.uint8)
    _check(_lib.bpx_synth_query(seed, index, length, _ptr(out)))
    return out[:length]


def synth_query_lengths(seed: int, n: int, kind: str, lo: float, hi: float, median=0.0, sigma=0.0):
    out = np.zeros(max(n, 1), dtype=np.uint32)
    _check(_lib.bpx_synth_query_lengths(seed, n, 0 if kind == "loguniform" else 1, lo, hi, median, sigma,
                                        _ptr(out)))
    return out[:n]


def sha256_hex(data: np.ndarray) -> str:
    a = np.ascontiguousarray(data)
    buf = C.create_string_buffer(65)
    _check(_lib.bpx_sha256_hex(_ptr(a), a.nbytes, buf))
    return buf.value.decode()


@dataclass
class SynthDB:
    residues: np.ndarray   # uint8, packed back to back
    offsets: np.ndarray    # uint64, num_subjects + 1
    lengths: np.ndarray    # uint32
    kinds: np.ndarray      # uint8: 0 background, 1 homolog, 2 biased
    sha256: str

    @property
    def num_subjects(self) -> int:
        return int(self.lengths.shape[0])

    @property
    def num_residues(self) -> int:
        return int(self.residues.shape[0])


def synth_db(spec: dict, queries=(), nthreads: int = 0) -> SynthDB:
    s = SynthSpec()
    for k, _ in SynthSpec._fields_:
        if k in spec:
            setattr(s, k, spec[k])
    qs = [np.ascontiguousarray(q, dtype=np.uint8) for q in queries]
    qptrs = (C.c_void_p * max(len(qs), 1))(*[q.ctypes.data for q in qs])
    qlens = np.array([len(q) for q in qs] or [0], dtype=np.uint32)
    h = _vp()
    _check(_lib.bpx_synth_db_generate(C.byref(s), qptrs, _ptr(qlens), len(qs), nthreads, C.byref(h)))
    try:
        res, off, ln, kd = _vp(), _vp(), _vp(), _vp()
        n, nres = C.c_uint32(), C.c_uint64()
        _check(_lib.bpx_synth_db_arrays(h, C.byref(res), C.byref(off), C.byref(ln), C.byref(n),
                                        C.byref(nres)))
        _check(_lib.bpx_synth_db_kinds(h, C.byref(kd)))
        n, nres = int(n.value), int(nres.value)

        def arr(p, count, dt):
            if count == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(np.ctypeslib.as_ctypes_type(dt))),
                                         shape=(count,)).copy()

        buf = C.create_string_buffer(65)
        _check(_lib.bpx_synth_db_sha256(h, buf))
        return SynthDB(arr(res, nres, np.uint8), arr(off, n + 1, np.uint64), arr(ln, n, np.uint32),
                       arr(kd, n, np.uint8), buf.value.decode())
    finally:
        _lib.bpx_synth_db_free(h)


# ---- device objects -----------------------------------------------------------

class Context:
    def __init__(self, device: int = 0):
        self._h = _vp()
        _check(_lib.bpx_context_create(device, C.byref(self._h)))

    def info(self):
        cus, lds, hbm = C.c_int32(), C.c_int64(), C.c_int64()
        _check(_lib.bpx_context_info(self._h, C.byref(cus), C.byref(lds), C.byref(hbm)))
        return {"num_cus": cus.value, "lds_bytes": lds.value, "hbm_bytes": hbm.value}

    def limit(self, hbm_bytes: int):
        """bpx_context_limit: cap the device bytes scans are planned with (0: the device's)."""
        _check(_lib.bpx_context_limit(self._h, hbm_bytes))

    def trim(self):
        """bpx_context_trim: free the buffers the context keeps cached for later scans."""
        _check(_lib.bpx_context_trim(self._h))

    def close(self):
        if self._h:
            _lib.bpx_context_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_db(self, residues, offsets, lengths) -> "Database":
        return Database(self, residues, offsets, lengths)

    def prepare_query(self, query, params: WordParams | None = None, matrix=None) -> "Query":
        return Query(self, query, params, matrix)

    def prepare_batch(self, queries, params: WordParams | None = None, matrix=None, x_dropoff=None,
                      cutoff=None, group_residues=0, group_queries=0, group_diag_slots=0) -> "Batch":
        return Batch(self, queries, params, matrix, x_dropoff, cutoff, group_residues, group_queries,
                     group_diag_slots)


class Database:
    def __init__(self, ctx: Context, residues, offsets, lengths):
        self.ctx = ctx
        r = np.ascontiguousarray(residues, dtype=np.uint8)
        o = np.ascontiguousarray(offsets, dtype=np.uint64)
        ln = np.ascontiguousarray(lengths, dtype=np.uint32)
        self.num_subjects = int(ln.shape[0])
        self._h = _vp()
        _check(_lib.bpx_db_load(ctx._h, _ptr(r), _ptr(o), _ptr(ln), self.num_subjects, C.byref(self._h)))
        n, nres, ab = C.c_uint32(), C.c_uint64(), C.c_uint64()
        _check(_lib.bpx_db_info(self._h, C.byref(n), C.byref(nres), C.byref(ab)))
        self.num_residues = int(nres.value)
        self.arena_bytes = int(ab.value)

    def close(self):
        if self._h:
            _lib.bpx_db_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Query:
    def __init__(self, ctx: Context, query, params: WordParams | None, matrix):
        self.ctx = ctx
        self.seq = np.ascontiguousarray(query, dtype=np.uint8)
        self.params = params if params is not None else default_params()
        m = None if matrix is None else np.ascontiguousarray(matrix, dtype=np.int8)
        self._h = _vp()
        _check(_lib.bpx_query_prepare(ctx._h, _ptr(self.seq), len(self.seq), _ptr(m),
                                      C.byref(self.params), C.byref(self._h)))

    def info(self):
        tb, lr, npw, ne = C.c_int64(), C.c_int32(), C.c_int32(), C.c_int64()
        _check(_lib.bpx_query_info(self._h, C.byref(tb), C.byref(lr), C.byref(npw), C.byref(ne)))
        return {"table_bytes": tb.value, "lds_resident": bool(lr.value), "present_words": npw.value,
                "entries": ne.value}

    def close(self):
        if self._h:
            _lib.bpx_query_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many queries staged together (packed into scan groups)."""

    def __init__(self, ctx: Context, queries, params, matrix, x_dropoff, cutoff, group_residues, group_queries,
                 group_diag_slots=0):
        self.ctx = ctx
        self.seqs = [np.ascontiguousarray(q, dtype=np.uint8) for q in queries]
        self.num_queries = len(self.seqs)
        self.params = params if params is not None else default_params()
        ptrs = (C.c_void_p * max(self.num_queries, 1))(*[q.ctypes.data for q in self.seqs])
        lens = np.array([len(q) for q in self.seqs] or [0], dtype=np.int32)
        m = None if matrix is None else np.ascontiguousarray(matrix, dtype=np.int8)
        xd = None if x_dropoff is None else np.ascontiguousarray(x_dropoff, dtype=np.int32)
        co = None if cutoff is None else np.ascontiguousarray(cutoff, dtype=np.int32)
        pol = BatchPolicy(group_residues, group_queries, group_diag_slots)
        self._h = _vp()
        _check(_lib.bpx_batch_prepare(ctx._h, self.num_queries, ptrs, _ptr(lens), _ptr(m), C.byref(self.params),
                                      _ptr(xd), _ptr(co), C.byref(pol), C.byref(self._h)))

    def info(self):
        nq, ng, tb = C.c_uint32(), C.c_uint32(), C.c_int64()
        _check(_lib.bpx_batch_info(self._h, C.byref(nq), C.byref(ng), C.byref(tb)))
        return {"queries": nq.value, "groups": ng.value, "table_bytes": tb.value}

    def close(self):
        if self._h:
            _lib.bpx_batch_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _launch(launch):
    if launch is None:
        return None
    if isinstance(launch, Launch):
        return launch
    return Launch(**launch)


class Result:
    """Records (ordered by query, subject, subject offset, query offset),
    their owner subjects and queries, per-(query, subject) counters (if
    requested) and per-query ungapped stats."""

    def __init__(self, handle, first, count, flags, num_queries=1):
        self._h = handle
        self.first, self.count, self.flags = first, count, flags
        self.num_queries = num_queries

    def wait(self) -> "Result":
        _check(_lib.bpx_result_wait(self._h))
        return self

    def timing(self):
        k, d = C.c_float(), C.c_float()
        _check(_lib.bpx_result_timing(self._h, C.byref(k), C.byref(d)))
        return {"kernel_ms": k.value, "device_ms": d.value}

    def summary(self):
        n, st = C.c_uint64(), UngappedStats()
        _check(_lib.bpx_result_summary(self._h, C.byref(n), C.byref(st)))
        return int(n.value), st.as_dict()

    def query_stats(self):
        arr = (UngappedStats * max(self.num_queries, 1))()
        _check(_lib.bpx_result_query_stats(self._h, arr, len(arr)))
        return [arr[i].as_dict() for i in range(self.num_queries)]

    def records(self, with_queries=False):
        n, _ = self.summary()
        rec = np.zeros(max(n, 1), dtype=HSP_DTYPE)
        sub = np.zeros(max(n, 1), dtype=np.uint32)
        qry = np.zeros(max(n, 1), dtype=np.uint32)
        _check(_lib.bpx_result_records(self._h, _ptr(rec), _ptr(sub), _ptr(qry), len(rec)))
        if with_queries:
            return rec[:n], sub[:n], qry[:n]
        return rec[:n], sub[:n]

    def subject_stats(self) -> np.ndarray:
        """Counters per (query, subject), shape (num_queries * count,) query-major."""
        n = self.num_queries * self.count
        st = np.zeros(max(n, 1), dtype=STATS_DTYPE)
        _check(_lib.bpx_result_subject_stats(self._h, _ptr(st), len(st)))
        return st[:n]

    def phases(self) -> dict:
        """bpx_result_phases (scan made with phases=True)."""
        st = PhaseStats()
        _check(_lib.bpx_result_phases(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def replay(self) -> "Result":
        """Re-runs a graph-mode scan (same buffers); call .wait() again."""
        _check(_lib.bpx_result_replay(self._h))
        return self

    def close(self):
        if self._h:
            _lib.bpx_result_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def scan(ctx: Context, query: Query, db: Database, first: int = 0, count: int | None = None,
         launch=None, subject_stats: bool = True, graph: bool = False, phases: bool = False) -> Result:
    """Asynchronous scan of subjects [first, first+count); call .wait().
    graph=True captures the device work in a HIP graph (Result.replay)."""
    if count is None:
        count = db.num_subjects - first
    flags = (OUT_SUBJECT_STATS if subject_stats else 0) | (OUT_GRAPH if graph else 0) | (OUT_PHASES if phases else 0)
    h = _vp()
    lp = _launch(launch)
    _check(_lib.bpx_scan(ctx._h, query._h, db._h, first, count, C.byref(lp) if lp is not None else None,
                         flags, C.byref(h)))
    return Result(h, first, count, flags)


def scan_batch(ctx: Context, batch: Batch, db: Database, first: int = 0, count: int | None = None,
               launch=None, subject_stats: bool = False, graph: bool = False, phases: bool = False) -> Result:
    """Asynchronous scan of every query of the batch in one persistent launch
    per launch class; graph=True captures it in a HIP graph (Result.replay)."""
    if count is None:
        count = db.num_subjects - first
    flags = (OUT_SUBJECT_STATS if subject_stats else 0) | (OUT_GRAPH if graph else 0) | (OUT_PHASES if phases else 0)
    h = _vp()
    lp = _launch(launch)
    _check(_lib.bpx_scan_batch(ctx._h, batch._h, db._h, first, count, C.byref(lp) if lp is not None else None,
                               flags, C.byref(h)))
    return Result(h, first, count, flags, batch.num_queries)


NOALLOC_MAX_SUBJECTS = 2097152
NOALLOC_SEG_RECORDS = 1048576


def scan_noalloc(ctx: Context, query: Query, db: Database, d_records: int, d_subjects: int, capacity: int,
                 d_stats: int | None = None):
    """bpx_scan_noalloc: a synchronous scan of the whole database that
    allocates nothing; the outputs are caller-owned device buffers given as
    addresses.  Returns (status, number of records, ungapped stats); a
    status other than 0 leaves the reason in last_error()."""
    n = C.c_uint64()
    st = UngappedStats()
    rc = _lib.bpx_scan_noalloc(ctx._h, query._h, db._h, d_records or None, d_subjects or None, capacity,
                               d_stats or None, C.byref(n), C.byref(st))
    return rc, int(n.value), st.as_dict()


def last_error() -> str:
    return _lib.bpx_last_error().decode(errors="replace")


def query_footprint(ctx: Context, db: Database, qlen: int, params: WordParams | None = None, launch=None):
    """bpx_query_footprint: (status, diagonal slots, waves, LDS bytes) of a
    qlen-residue query scanned alone against db."""
    d, w, lds = C.c_uint32(), C.c_int32(), C.c_size_t()
    lp = _launch(launch)
    rc = _lib.bpx_query_footprint(ctx._h, db._h, qlen, C.byref(params) if params is not None else None,
                                  C.byref(lp) if lp is not None else None, C.byref(d), C.byref(w), C.byref(lds))
    return rc, int(d.value), int(w.value), int(lds.value)


def check_memory(ctx: Context, query: Query, db: Database, first=0, count=None, launch=None):
    if count is None:
        count = db.num_subjects - first
    dev, lds = C.c_size_t(), C.c_size_t()
    lp = _launch(launch)
    rc = _lib.bpx_check_memory(ctx._h, db._h, query._h, first, count,
                               C.byref(lp) if lp is not None else None, C.byref(dev), C.byref(lds))
    return rc, int(dev.value), int(lds.value)


def _struct_dict(st):
    return {k: getattr(st, k) for k, _ in st._fields_}


def autotune_batch(ctx: Context, queries, db: Database, sample_first: int = 0, sample_count: int | None = None,
                   params: WordParams | None = None, matrix=None, x_dropoff=None, cutoff=None):
    """Times candidate packing policies / launch shapes on subjects
    [sample_first, sample_first + sample_count) and returns
    (policy dict -- group_residues, group_queries, group_diag_slots, the
    keyword arguments of Context.prepare_batch -- launch dict, trials)."""
    seqs = [np.ascontiguousarray(q, dtype=np.uint8) for q in queries]
    if sample_count is None:
        sample_count = max(1, db.num_subjects // 20)
    ptrs = (C.c_void_p * max(len(seqs), 1))(*[q.ctypes.data for q in seqs])
    lens = np.array([len(q) for q in seqs] or [0], dtype=np.int32)
    prm = params if params is not None else default_params()
    m = None if matrix is None else np.ascontiguousarray(matrix, dtype=np.int8)
    xd = None if x_dropoff is None else np.ascontiguousarray(x_dropoff, dtype=np.int32)
    co = None if cutoff is None else np.ascontiguousarray(cutoff, dtype=np.int32)
    out = TuneResult()
    trials = (TuneTrial * 64)()
    nt = C.c_uint32()
    _check(_lib.bpx_autotune_batch(ctx._h, len(seqs), ptrs, _ptr(lens), _ptr(m), C.byref(prm), _ptr(xd), _ptr(co),
                                   db._h, sample_first, sample_count, C.byref(out), trials, 64, C.byref(nt)))
    tr = [{"policy": _struct_dict(t.policy), "launch": _struct_dict(t.launch), "sample_ms": round(float(t.sample_ms), 3)}
          for t in trials[:min(nt.value, 64)]]
    return _struct_dict(out.policy), _struct_dict(out.launch), tr


CM_FEATURES = 32
CM_PARAMS = 7312


def word_counts(queries, threshold=11, word_length=3, matrix=None) -> np.ndarray:
    """Sum over the queries of each word's lookup-cell size (the number of
    hits a subject word would make against all of them), saturated to u16."""
    tot = np.zeros(1 << (5 * word_length), dtype=np.int64)
    for q in queries:
        cs, _ = table_build(q, threshold, word_length, matrix)
        tot += np.diff(cs.astype(np.int64))
    return np.minimum(tot, 65535).astype(np.uint16)


def db_features(ctx: Context, db: Database, counts: np.ndarray, word_length=3) -> np.ndarray:
    out = np.zeros((max(db.num_subjects, 1), CM_FEATURES), dtype=np.float32)
    wc = np.ascontiguousarray(counts, dtype=np.uint16)
    _check(_lib.bpx_db_features(ctx._h, db._h, _ptr(wc), word_length, _ptr(out)))
    return out[:db.num_subjects]


def db_set_cost(db: Database, cost=None):
    """Order (and bin) the database's scan work by descending cost; None
    restores the longest-first default."""
    c = None if cost is None else np.ascontiguousarray(cost, dtype=np.float32)
    _check(_lib.bpx_db_set_cost(db._h, _ptr(c)))


class CostModel:
    """32-64-64-1 MLP (bf16 weights/activations, fp32 accumulation) trained
    and run on the GPU by the library's MFMA kernels."""

    def __init__(self, device=0, seed=0):
        self._h = _vp()
        rc = _lib.bpx_costmodel_create(device, seed, C.byref(self._h))
        if rc:
            raise BpxError(rc, _lib.bpx_costmodel_last_error().decode())

    def _chk(self, rc):
        if rc:
            raise BpxError(rc, _lib.bpx_costmodel_last_error().decode())

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(CM_PARAMS, dtype=np.float32)
        self._chk(_lib.bpx_costmodel_params(self._h, _ptr(p), 0))
        return p

    @params.setter
    def params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.shape == (CM_PARAMS,)
        self._chk(_lib.bpx_costmodel_params(self._h, _ptr(a), 1))

    def train(self, X, y, epochs=8, batch=4096, lr=3e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        losses = np.zeros(epochs, dtype=np.float32)
        self._chk(_lib.bpx_costmodel_train(self._h, _ptr(X), _ptr(y), len(y), epochs, batch, lr, seed,
                                           _ptr(losses)))
        return losses

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros(max(len(X), 1), dtype=np.float32)
        self._chk(_lib.bpx_costmodel_predict(self._h, _ptr(X), len(X), _ptr(out)))
        return out[:len(X)]

    def grad(self, X, y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(CM_PARAMS, dtype=np.float32)
        loss = C.c_float()
        self._chk(_lib.bpx_costmodel_grad(self._h, _ptr(X), _ptr(y), len(y), C.byref(loss), _ptr(g)))
        return float(loss.value), g

    def close(self):
        if self._h:
            _lib.bpx_costmodel_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching query server over a resident database (include/bpx.h
    bpx_server_*): submit() queues a query, wait() blocks for it, records()
    fetches its records.  Each result equals scanning the query alone."""

    def __init__(self, ctx: Context, db: Database, max_batch=0, max_inflight=0, max_wait_us=0, hbm_budget=0,
                 group_residues=0, group_queries=0, launch=None, params: WordParams | None = None, diag_budget=0,
                 group_diag_slots=0):
        cfg = ServerConfig()
        cfg.max_batch, cfg.max_inflight, cfg.max_wait_us, cfg.hbm_budget = max_batch, max_inflight, max_wait_us, hbm_budget
        cfg.diag_budget = diag_budget
        if params is not None:
            cfg.params = params
        cfg.policy = BatchPolicy(group_residues, group_queries, group_diag_slots)
        if launch is not None:
            cfg.launch = _launch(launch)
        self.ctx, self.db = ctx, db
        self._h = _vp()
        _check(_lib.bpx_server_create(ctx._h, db._h, C.byref(cfg), C.byref(self._h)))

    def submit(self, query) -> int:
        q = np.ascontiguousarray(query, dtype=np.uint8)
        t = C.c_uint64()
        _check(_lib.bpx_server_submit(self._h, _ptr(q), len(q), C.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> dict:
        info = RequestInfo()
        _check(_lib.bpx_server_wait(self._h, ticket, timeout_ms, C.byref(info)))
        return {"num_records": int(info.num_records), "stats": info.stats.as_dict(),
                "latency_ms": float(info.latency_ms), "queue_ms": float(info.queue_ms),
                "batch_size": int(info.batch_size)}

    def records(self, ticket: int, n: int):
        rec = np.zeros(max(n, 1), dtype=HSP_DTYPE)
        sub = np.zeros(max(n, 1), dtype=np.uint32)
        _check(_lib.bpx_server_records(self._h, ticket, _ptr(rec), _ptr(sub), len(rec)))
        return rec[:n], sub[:n]

    def counters(self):
        b, c = C.c_uint64(), C.c_uint64()
        _check(_lib.bpx_server_counters(self._h, C.byref(b), C.byref(c)))
        return {"batches": b.value, "completed": c.value}

    def close(self):
        if self._h:
            _lib.bpx_server_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _vchk(rc):
    if rc:
        raise BpxError(rc, _lib.bpx_volume_last_error().decode())


def volume_write(base: str, residues, offsets, lengths, title: str = ""):
    """Writes an NCBI version-4 protein volume base.pin / .psq / .phr."""
    r = np.ascontiguousarray(residues, dtype=np.uint8)
    o = np.ascontiguousarray(offsets, dtype=np.uint64)
    ln = np.ascontiguousarray(lengths, dtype=np.uint32)
    _vchk(_lib.bpx_volume_write(base.encode(), title.encode(), _ptr(r), _ptr(o), _ptr(ln), len(ln)))


def _volume_arrays(h):
    try:
        n, tot, mx, title = C.c_uint32(), C.c_uint64(), C.c_uint32(), C.c_char_p()
        _vchk(_lib.bpx_volume_info(h, C.byref(n), C.byref(tot), C.byref(mx), C.byref(title)))
        rp, op, lp = _vp(), _vp(), _vp()
        _vchk(_lib.bpx_volume_arrays(h, C.byref(rp), C.byref(op), C.byref(lp)))
        n, tot = int(n.value), int(tot.value)

        def arr(p, count, dt):
            if count == 0 or not p.value:
                return np.zeros(count, dtype=dt)
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(count,)).copy()

        return arr(rp, tot, np.uint8), arr(op, n + 1, np.uint64), arr(lp, n, np.uint32), title.value.decode()
    finally:
        _lib.bpx_volume_destroy(h)


def volume_read(base: str):
    """Reads base.pin / base.psq: (residues, offsets[n+1], lengths[n], title)."""
    h = _vp()
    _vchk(_lib.bpx_volume_read(base.encode(), C.byref(h)))
    return _volume_arrays(h)


def gpu_info(gpuinfo: str):
    """The header of a .gpuinfo file: (volumes, group size, percentage)."""
    nv, gs, pc = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _vchk(_lib.bpx_gpu_info(gpuinfo.encode(), C.byref(nv), C.byref(gs), C.byref(pc)))
    return int(nv.value), int(gs.value), int(pc.value)


def gpu_volume_read(gpuinfo: str, volume: int = 0):
    """Volume `volume` of a .gpu database, de-interleaved:
    (residues, offsets[n+1], lengths[n], "")."""
    h = _vp()
    _vchk(_lib.bpx_gpu_volume_read(gpuinfo.encode(), volume, C.byref(h)))
    return _volume_arrays(h)


def db_volumes(base: str):
    need, cnt = C.c_size_t(), C.c_uint32()
    _vchk(_lib.bpx_db_volumes(base.encode(), None, 0, C.byref(need), C.byref(cnt)))
    buf = C.create_string_buffer(need.value)
    _vchk(_lib.bpx_db_volumes(base.encode(), buf, need.value, None, None))
    return buf.value.decode().split("\n") if cnt.value else []


def load_volumes(ctx: "Context", base: str) -> "Database":
    """Every volume of an NCBI protein database (alias or single volume) as one
    resident arena; subject i is OID i."""
    db = Database.__new__(Database)
    db.ctx = ctx
    db._h = _vp()
    rc = _lib.bpx_db_load_volumes(ctx._h, base.encode(), C.byref(db._h))
    if rc:
        msg = _lib.bpx_volume_last_error().decode() or _lib.bpx_last_error().decode()
        raise BpxError(rc, msg)
    n, nres, ab = C.c_uint32(), C.c_uint64(), C.c_uint64()
    _check(_lib.bpx_db_info(db._h, C.byref(n), C.byref(nres), C.byref(ab)))
    db.num_subjects, db.num_residues, db.arena_bytes = int(n.value), int(nres.value), int(ab.value)
    return db

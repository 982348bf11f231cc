"""The allocation-free scan (bpx_scan_noalloc) against bpx_scan, through the
C ABI: two calls in a row, the second on a smaller database with a shorter
query, each bit-equal to bpx_scan -- records, subject ids, per-subject
counters and ungapped stats.  A record capacity below the record count is
refused with the count reported, and a database above the subject limit is
refused before any device work."""
import ctypes as C

import numpy as np
import pytest

import bpx
from bpx import workloads as W

pytestmark = pytest.mark.gpu

# The HIP runtime libbpx.so is linked against (already loaded by it, found by
# its soname): the test allocates its output buffers there, as a caller of the
# C ABI would.
_hip = C.CDLL("libamdhip64.so.7")
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
_hip.hipDeviceSynchronize.argtypes = []
D2H = 2
E_ARG, E_BUDGET = -1, -5


class DevBuf:
    """Caller-owned device memory, filled with 0xFF so that stale bytes show."""

    def __init__(self, nbytes):
        self.p = C.c_void_p()
        self.nbytes = max(int(nbytes), 1)
        assert _hip.hipMalloc(C.byref(self.p), self.nbytes) == 0
        assert _hip.hipMemset(self.p, 0xFF, self.nbytes) == 0
        assert _hip.hipDeviceSynchronize() == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        if count:
            assert _hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    def close(self):
        _hip.hipFree(self.p)


@pytest.fixture(scope="module")
def ctx():
    assert bpx.available(), "no HIP device visible"
    c = bpx.Context(0)
    yield c
    c.close()


def with_bpx_scan(ctx, dq, ddb):
    r = bpx.scan(ctx, dq, ddb, subject_stats=True).wait()
    recs, subs = r.records()
    st = r.subject_stats()
    ust = r.summary()[1]
    r.close()
    return recs, subs, st, ust


def noalloc(ctx, dq, ddb, n_sub, capacity):
    recs = DevBuf(capacity * 24)
    subs = DevBuf(capacity * 4)
    stats = DevBuf(n_sub * 12)
    rc, n, ust = bpx.scan_noalloc(ctx, dq, ddb, recs.p.value, subs.p.value, capacity, stats.p.value)
    k = min(n, capacity)
    out = (rc, n, recs.get(np.int32, 6 * k).reshape(-1, 6), subs.get(np.uint32, k),
           stats.get(np.int32, 3 * n_sub).reshape(-1, 3), ust)
    for b in (recs, subs, stats):
        b.close()
    return out


def assert_equal_to_bpx_scan(ctx, q, db):
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    dq = ctx.prepare_query(q)
    recs, subs, st, ust = with_bpx_scan(ctx, dq, ddb)
    assert len(recs) > 0
    rc, n, grecs, gsubs, gst, gust = noalloc(ctx, dq, ddb, db.num_subjects, len(recs))
    assert rc == 0, bpx.last_error()
    assert n == len(recs)
    np.testing.assert_array_equal(grecs, recs.view(np.int32).reshape(-1, 6))
    np.testing.assert_array_equal(gsubs, subs)
    np.testing.assert_array_equal(gst, st.view(np.int32).reshape(-1, 3))
    assert gust == ust
    # one record short: refused, the count reported, nothing past capacity written
    rc, n, grecs, gsubs, _, _ = noalloc(ctx, dq, ddb, db.num_subjects, len(recs) - 1)
    assert rc == E_ARG and n == len(recs), bpx.last_error()
    dq.close()
    ddb.close()
    return len(recs)


def test_noalloc_equals_bpx_scan_twice_then_smaller(ctx):
    q1, db1 = W.config1(16)
    n1 = assert_equal_to_bpx_scan(ctx, q1, db1)
    # the second call: a smaller database and a shorter query, on the same
    # scratch, which still holds the first call's contents
    q2 = bpx.synth_query(71, 0, 180)
    spec = dict(seed=72, num_subjects=2500, len_median=250, len_sigma=0.6, len_min=30, len_max=3000,
                homolog_count=25, identity_min=0.5, identity_max=0.7, indel_min=0.0, indel_max=0.02)
    db2 = bpx.synth_db(spec, [q2], nthreads=16)
    n2 = assert_equal_to_bpx_scan(ctx, q2, db2)
    print(f"allocation-free scans equal bpx_scan: {n1} then {n2} records", flush=True)


def test_noalloc_refuses_a_database_above_its_limit(ctx):
    n = bpx.NOALLOC_MAX_SUBJECTS + 1
    lens = np.ones(n, dtype=np.uint32)
    res = np.ones(n, dtype=np.uint8)
    offs = np.arange(n, dtype=np.uint64)
    ddb = ctx.load_db(res, offs, lens)
    dq = ctx.prepare_query(bpx.synth_query(73, 0, 50))
    rc, num, _ = bpx.scan_noalloc(ctx, dq, ddb, 0, 0, 0, None)
    assert rc == E_BUDGET and num == 0
    assert "limit" in bpx.last_error()
    dq.close()
    ddb.close()

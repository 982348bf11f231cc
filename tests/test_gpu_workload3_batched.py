"""GPU parity at the full size of the batched workload (1,024 queries of
log-uniform length 50-2,000 x 500,000 subjects): the autotuned batch,
captured as one HIP graph and replayed, equals the CPU oracle for every 16th
query (the whole database each) and the per-query HIP scan for all 1,024
queries -- records, subject and query ids, per-subject counters and
per-query ungapped stats."""
import time

import numpy as np
import pytest

import bpx
from bpx import workloads as W
from fullsize import THREADS, assert_same_as_oracle, assert_same_snapshot, query_slices, say, snapshot

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    assert bpx.available(), "no HIP device visible"
    c = bpx.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def batched(ctx):
    t = time.perf_counter()
    qs, db = W.config3(THREADS)
    assert db.sha256 == W.SHA256[3]
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    policy, launch, trials = bpx.autotune_batch(ctx, qs, ddb, 0, db.num_subjects // 10)
    say(f"batched workload: autotuned {policy} launch={launch} ({time.perf_counter() - t:.1f} s)")
    b = ctx.prepare_batch(qs, **policy)
    g = bpx.scan_batch(ctx, b, ddb, launch=launch, subject_stats=True, graph=True).wait()
    snap = snapshot(g, len(qs), db.num_subjects)
    say(f"batched workload: {len(snap['recs'])} records ({time.perf_counter() - t:.1f} s)")
    # a replay of the graph gives the same result
    g.replay().wait()
    assert_same_snapshot(snapshot(g, len(qs), db.num_subjects), snap)
    g.close()
    b.close()
    say(f"batched workload: replay equal ({time.perf_counter() - t:.1f} s)")
    yield qs, db, ddb, snap
    ddb.close()


@pytest.mark.timeout(900)
def test_batched_workload_equals_oracle_every_16th_query(batched):
    qs, db, _, snap = batched
    sl = query_slices(snap["qids"], len(qs))
    t = time.perf_counter()
    n = 0
    for i in range(0, len(qs), 16):
        a, b = sl[i]
        n += assert_same_as_oracle(qs[i], db, snap["recs"][a:b], snap["subs"][a:b], snap["st"][i], snap["qstats"][i])
        if i % 256 == 0:
            say(f"  query {i} ({len(qs[i])} aa): equal to the oracle ({time.perf_counter() - t:.1f} s)")
    say(f"64 queries equal to the oracle, {n} records ({time.perf_counter() - t:.1f} s)")


@pytest.mark.timeout(900)
def test_batched_workload_equals_per_query_scans(ctx, batched):
    qs, db, ddb, snap = batched
    sl = query_slices(snap["qids"], len(qs))
    t = time.perf_counter()
    for i, q in enumerate(qs):
        dq = ctx.prepare_query(q)
        r = bpx.scan(ctx, dq, ddb, subject_stats=True).wait()
        recs, subs = r.records()
        st = r.subject_stats()
        _, ust = r.summary()
        r.close()
        dq.close()
        a, b = sl[i]
        np.testing.assert_array_equal(recs.view(np.int32), snap["recs"][a:b].view(np.int32))
        np.testing.assert_array_equal(subs, snap["subs"][a:b])
        np.testing.assert_array_equal(st.view(np.int32), snap["st"][i].view(np.int32))
        assert ust == snap["qstats"][i], i
        if i % 256 == 0:
            say(f"  query {i}: batch row equals its own scan ({time.perf_counter() - t:.1f} s)")
    say(f"all {len(qs)} batch rows equal the per-query scans ({time.perf_counter() - t:.1f} s)")

"""GPU parity of the server at full size: a Poisson burst of single queries
against the resident 2,000,000-subject database; every served result equals
the plain scan of the same query, and every 4th equals the CPU oracle."""
import time

import numpy as np
import pytest

import bpx
from bpx import workloads as W
from fullsize import THREADS, assert_same_as_oracle, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    assert bpx.available(), "no HIP device visible"
    c = bpx.Context(0)
    yield c
    c.close()


@pytest.mark.timeout(900)
def test_serving_burst_equals_plain_scans_and_oracle(ctx):
    t = time.perf_counter()
    q0, db = W.config2(THREADS)
    assert db.sha256 == W.SHA256[2]
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    qs = W.serving_queries(48, seed=55)
    srv = bpx.Server(ctx, ddb)
    rng = np.random.default_rng(56)
    tickets = []
    for q in qs:  # Poisson arrivals at 200 queries/s: batches form while others run
        tickets.append(srv.submit(q))
        time.sleep(float(rng.exponential(1.0 / 200.0)))
    served = []
    for tk in tickets:
        info = srv.wait(tk)
        recs, subs = srv.records(tk, info["num_records"])
        served.append((recs.view(np.int32).copy(), subs.copy(), info["stats"], info["batch_size"]))
    c = srv.counters()
    srv.close()
    sizes = [s[3] for s in served]
    say(f"serving: {len(qs)} requests in {c['batches']} batches (sizes {sorted(set(sizes))}) "
        f"({time.perf_counter() - t:.1f} s)")
    assert max(sizes) > 1
    for i, q in enumerate(qs):
        dq = ctx.prepare_query(q)
        r = bpx.scan(ctx, dq, ddb, subject_stats=(i % 4 == 0)).wait()
        recs, subs = r.records()
        ust = r.summary()[1]
        np.testing.assert_array_equal(served[i][0], recs.view(np.int32))
        np.testing.assert_array_equal(served[i][1], subs)
        assert served[i][2] == ust, i
        if i % 4 == 0:
            assert_same_as_oracle(q, db, recs, subs, r.subject_stats(), ust)
            say(f"  request {i} ({len(q)} aa): served = plain = oracle ({time.perf_counter() - t:.1f} s)")
        r.close()
        dq.close()
    ddb.close()

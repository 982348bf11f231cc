"""Serving: queries submitted to the continuous-batching server return,
each, exactly what scanning it alone returns -- alone on the server, in a
mix of concurrent arrivals, and against the plain single-query scan."""
import sys
import time

import numpy as np
import pytest

import bpx
import pyoracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    assert bpx.available(), "no HIP device visible"
    c = bpx.Context(0)
    yield c
    c.close()


def test_results_same_alone_and_in_the_mix(ctx):
    lens = bpx.synth_query_lengths(50, 40, "lognormal", 20, 3000, 350, 0.6)
    qs = [bpx.synth_query(50, i, int(n)) for i, n in enumerate(lens)]
    spec = dict(seed=51, num_subjects=20000, len_median=300, len_sigma=0.6, len_min=30, len_max=5000,
                homolog_fraction=0.02, identity_min=0.3, identity_max=0.7, indel_min=0.0, indel_max=0.03)
    sdb = bpx.synth_db(spec, qs, nthreads=16)
    db = ctx.load_db(sdb.residues, sdb.offsets[:-1], sdb.lengths)

    def plain(q):
        dq = ctx.prepare_query(q)
        r = bpx.scan(ctx, dq, db, subject_stats=False).wait()
        recs, subs = r.records()
        st = r.summary()[1]
        r.close()
        dq.close()
        return recs.view(np.int32).copy(), subs.copy(), st

    want = [plain(q) for q in qs]
    srv = bpx.Server(ctx, db, max_batch=16, max_wait_us=3000)

    def fetch(t):
        info = srv.wait(t)
        recs, subs = srv.records(t, info["num_records"])
        return info, (recs.view(np.int32).copy(), subs.copy(), info["stats"])

    # alone: one request at a time
    for i in (0, 7, 19):
        info, got = fetch(srv.submit(qs[i]))
        assert info["batch_size"] == 1
        np.testing.assert_array_equal(got[0], want[i][0])
        np.testing.assert_array_equal(got[1], want[i][1])
        assert got[2] == want[i][2]
    # the mix: arrivals spread over a few milliseconds, several batches in flight
    rng = np.random.default_rng(52)
    tickets = []
    for q in qs:
        tickets.append(srv.submit(q))
        time.sleep(float(rng.exponential(0.0005)))
    sizes = []
    for i, t in enumerate(tickets):
        info, got = fetch(t)
        sizes.append(info["batch_size"])
        np.testing.assert_array_equal(got[0], want[i][0])
        np.testing.assert_array_equal(got[1], want[i][1])
        assert got[2] == want[i][2], i
    c = srv.counters()
    say(f"batch sizes {sizes}; counters {c}")
    assert max(sizes) > 1 and c["completed"] == len(qs) + 3
    srv.close()
    db.close()


def test_bad_requests_are_refused(ctx):
    sdb = bpx.synth_db(dict(seed=53, num_subjects=100, len_median=200, len_sigma=0.5, len_min=30, len_max=500))
    db = ctx.load_db(sdb.residues, sdb.offsets[:-1], sdb.lengths)
    srv = bpx.Server(ctx, db)
    with pytest.raises(bpx.BpxError):
        srv.submit(np.full(10, 40, dtype=np.uint8))  # residue code >= 28
    with pytest.raises(bpx.BpxError):
        srv.wait(12345, 10)  # unknown ticket
    srv.close()
    db.close()


def test_diagonal_budget_splits_batches_and_results_equal_the_oracle(ctx):
    """Admission by the diagonal-table budget: with room for four 512-slot
    queries per batch, a burst of 24 queries is served in batches of at most
    four, and every result equals the oracle.  A query too long to run at
    all is refused at submit, and the others are not affected."""
    qs = [bpx.synth_query(57, i, 300 + 6 * i) for i in range(24)]  # 300-438 aa: 512 diagonal slots each
    spec = dict(seed=58, num_subjects=6000, len_median=300, len_sigma=0.6, len_min=30, len_max=3000,
                homolog_fraction=0.03, identity_min=0.4, identity_max=0.7, indel_min=0.0, indel_max=0.02)
    sdb = bpx.synth_db(spec, qs, nthreads=16)
    db = ctx.load_db(sdb.residues, sdb.offsets[:-1], sdb.lengths)
    srv = bpx.Server(ctx, db, diag_budget=4 * 512, max_wait_us=20000)
    with pytest.raises(bpx.BpxError, match="diagonal slots"):
        srv.submit(np.full(65500, 1, dtype=np.uint8))
    tickets = [srv.submit(q) for q in qs]
    sizes = []
    for q, t in zip(qs, tickets):
        info = srv.wait(t)
        recs, subs = srv.records(t, info["num_records"])
        sizes.append(info["batch_size"])
        counters, recoff, orecs = O.run_db(q, sdb.residues, sdb.offsets, sdb.lengths, None, 16)
        assert len(recs) == len(orecs)
        np.testing.assert_array_equal(recs.view(np.int32).reshape(-1, 6), orecs.view(np.int32).reshape(-1, 6))
        np.testing.assert_array_equal(subs, np.repeat(np.arange(sdb.num_subjects, dtype=np.uint32),
                                                      np.diff(recoff.astype(np.int64))))
        has = counters[:, 0] > 0
        assert info["stats"]["lookup_hits"] == int(counters[has, 0].sum())
        assert info["stats"]["init_extends"] == int(counters[has, 2].sum())
    c = srv.counters()
    say(f"batch sizes {sizes}; counters {c}")
    assert max(sizes) <= 4 and c["batches"] >= 6 and c["completed"] == len(qs)
    srv.close()
    db.close()

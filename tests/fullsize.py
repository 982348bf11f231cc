"""Helpers of the full-size workload tests: oracle comparison of one query's
slice of a batched result, and snapshots of batched results."""
import sys

import numpy as np
import pyoracle as O

THREADS = 16


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def query_slices(qids, nq):
    """Record index range of each query (records are ordered by query first)."""
    bounds = np.searchsorted(qids, np.arange(nq + 1, dtype=np.int64))
    return [(int(bounds[i]), int(bounds[i + 1])) for i in range(nq)]


def assert_same_as_oracle(q, db, recs, subs, st, qstats):
    counters, recoff, orecs = O.run_db(q, db.residues, db.offsets, db.lengths, None, THREADS)
    g = np.stack([st["hits"], st["successful_extensions"], st["hits_extended"]], axis=1)
    np.testing.assert_array_equal(g, counters)
    assert len(recs) == len(orecs)
    np.testing.assert_array_equal(recs.view(np.int32).reshape(-1, 6), orecs.view(np.int32).reshape(-1, 6))
    want_sub = np.repeat(np.arange(db.num_subjects, dtype=np.uint32), np.diff(recoff.astype(np.int64)))
    np.testing.assert_array_equal(subs, want_sub)
    has = counters[:, 0] > 0
    assert qstats["lookup_hits"] == int(counters[has, 0].sum())
    assert qstats["num_seqs_lookup_hits"] == int(has.sum())
    assert qstats["init_extends"] == int(counters[has, 2].sum())
    assert qstats["good_init_extends"] == int(counters[has, 1].sum())
    assert qstats["num_seqs_passed"] == int((counters[has, 1] > 0).sum())
    return len(orecs)


def snapshot(r, nq, nsub):
    recs, subs, qids = r.records(with_queries=True)
    return dict(recs=recs, subs=subs, qids=qids, st=r.subject_stats().reshape(nq, nsub),
                qstats=r.query_stats())


def assert_same_snapshot(a, b):
    for k in ("recs", "subs", "qids"):
        np.testing.assert_array_equal(a[k].view(np.int32) if k == "recs" else a[k],
                                      b[k].view(np.int32) if k == "recs" else b[k])
    np.testing.assert_array_equal(a["st"].view(np.int32), b["st"].view(np.int32))
    assert a["qstats"] == b["qstats"]



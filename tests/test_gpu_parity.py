"""GPU parity: the HIP word finder (through the C ABI) against the CPU oracle.

Bit-exact on every integer: per-subject counters {hits,
successful_extensions, hits_extended} and every record field, in the same
order (subject, then subject offset, then query offset -- the sequential
emission order).  Workloads: the hand-derived known answers, naive
straight-line cases, diagonal aliasing, edge lengths, both lookup-table
placements (LDS and HBM), word length 2, forced record-segment overflow, and
the full 10,000-subject and 2,000,000-subject single-query databases.
"""
import json
import os
import sys

import numpy as np
import pytest

import bpx
import bpx_letters
import pyoracle as O
import ref_straight as R

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STD20 = np.array([1, 16, 13, 4, 3, 15, 5, 7, 8, 9, 11, 10, 12, 6, 14, 17, 18, 20, 22, 19], dtype=np.uint8)
THREADS = 16


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    assert bpx.available(), "no HIP device visible"
    c = bpx.Context(0)
    say(f"device: {c.info()}")
    yield c
    c.close()


def pack(subjects):
    lens = np.array([len(s) for s in subjects], dtype=np.uint32)
    offs = np.concatenate([[0], np.cumsum(lens.astype(np.uint64))]).astype(np.uint64)
    res = np.concatenate(subjects).astype(np.uint8) if len(subjects) and lens.sum() else np.zeros(0, np.uint8)
    return res, offs, lens


def gpu_run(ctx, query, res, offs, lens, params=None, matrix=None, first=0, count=None, launch=None):
    db = ctx.load_db(res, offs[:-1] if len(offs) == len(lens) + 1 else offs, lens)
    q = ctx.prepare_query(query, params, matrix)
    r = bpx.scan(ctx, q, db, first, count, launch=launch).wait()
    recs, subs = r.records()
    st = r.subject_stats()
    n, ustats = r.summary()
    info = q.info()
    r.close()
    q.close()
    db.close()
    return st, recs, subs, ustats, info


def oracle_run(query, res, offs, lens, params=None, matrix=None):
    prm = None
    if params is not None:
        prm = O.params(len(query), params.word_length, params.threshold, params.window, params.x_dropoff,
                       params.cutoff_score, params.diag_mask)
    return O.run_db(query, res, offs, lens, prm, THREADS, matrix)


def assert_parity(gpu, orc, first=0):
    st, recs, subs, ustats, _ = gpu
    counters, recoff, orecs = orc
    n = len(st)
    g = np.stack([st["hits"], st["successful_extensions"], st["hits_extended"]], axis=1)
    np.testing.assert_array_equal(g, counters)
    assert len(recs) == len(orecs)
    np.testing.assert_array_equal(recs.view(np.int32).reshape(-1, 6), orecs.view(np.int32).reshape(-1, 6))
    want_sub = np.repeat(np.arange(first, first + n, dtype=np.uint32),
                         np.diff(recoff.astype(np.int64)).astype(np.int64))
    np.testing.assert_array_equal(subs, want_sub)
    # query-level ungapped stats (Blast_UngappedStatsUpdate sums)
    has = counters[:, 0] > 0
    assert ustats["lookup_hits"] == int(counters[has, 0].sum())
    assert ustats["num_seqs_lookup_hits"] == int(has.sum())
    assert ustats["init_extends"] == int(counters[has, 2].sum())
    assert ustats["good_init_extends"] == int(counters[has, 1].sum())
    assert ustats["num_seqs_passed"] == int((counters[has, 1] > 0).sum())


def identity_matrix():
    M = np.full((28, 28), -4, dtype=np.int8)
    np.fill_diagonal(M, 5)
    M[0, :] = -128
    M[:, 0] = -128
    return M


def test_known_answers(ctx):
    with open(os.path.join(ROOT, "tests", "golden", "known_answers.json")) as f:
        cases = json.load(f)["cases"]
    M = identity_matrix()
    for case in cases:
        say(f"known answer: {case['name']}")
        q = bpx_letters.encode(case["query"])
        s = bpx_letters.encode(case["subject"])
        p = bpx.default_params(**{("cutoff_score" if k == "cutoff" else k): v
                                  for k, v in case.get("params", {}).items()})
        res, offs, lens = pack([s])
        st, recs, subs, _, _ = gpu_run(ctx, q, res, offs, lens, p, M)
        assert [int(st[0]["hits"]), int(st[0]["successful_extensions"]), int(st[0]["hits_extended"])] == \
            case["counters"], case["name"]
        assert recs.view(np.int32).reshape(-1, 6).tolist() == case["records"], case["name"]


def test_straightline_small(ctx):
    rng = np.random.default_rng(77)
    M = O.blosum62()
    q = rng.choice(STD20, 45)
    subjects = []
    for i in range(60):
        s = rng.choice(STD20, int(rng.integers(0, 150)))
        if i % 2:
            seg = q[int(rng.integers(0, 30)):].copy()
            mut = rng.random(len(seg)) < 0.3
            seg[mut] = rng.choice(STD20, int(mut.sum()))
            s = np.concatenate([s, seg])
        subjects.append(s)
    res, offs, lens = pack(subjects)
    p = bpx.default_params(cutoff_score=20)
    st, recs, subs, _, _ = gpu_run(ctx, q, res, offs, lens, p)
    k = 0
    for i, s in enumerate(subjects):
        h, sv, ex, rr = R.scan(q, s, M, cutoff=20)
        assert [int(st[i]["hits"]), int(st[i]["successful_extensions"]), int(st[i]["hits_extended"])] == \
            [h, sv, ex], i
        got = [tuple(int(x) for x in r) for r in recs[k:k + sv].view(np.int32).reshape(-1, 6)]
        assert sorted(got) == sorted(rr), i
        k += sv


@pytest.mark.parametrize("diag_len,window", [(4, 5), (8, 40), (16, 12), (64, 40)])
def test_diagonal_aliasing(ctx, diag_len, window):
    say(f"aliasing diag_len={diag_len} window={window}")
    rng = np.random.default_rng(diag_len * 100 + window)
    q = rng.choice(STD20[:5], 60)
    subjects = [rng.choice(STD20[:5], int(rng.integers(1, 400))) for _ in range(300)]
    res, offs, lens = pack(subjects)
    p = bpx.default_params(window=window, diag_mask=diag_len - 1, cutoff_score=25)
    assert_parity(gpu_run(ctx, q, res, offs, lens, p), oracle_run(q, res, offs, lens, p))


def test_edge_lengths_and_subrange(ctx):
    rng = np.random.default_rng(5)
    q = rng.choice(STD20, 120)
    subjects = [np.zeros(0, np.uint8), rng.choice(STD20, 1), rng.choice(STD20, 2), rng.choice(STD20, 3)]
    subjects += [np.concatenate([rng.choice(STD20, int(rng.integers(0, 300))), q]) for _ in range(500)]
    res, offs, lens = pack(subjects)
    assert_parity(gpu_run(ctx, q, res, offs, lens), oracle_run(q, res, offs, lens))
    # a sub-range of the volume, as the reference hands the GPU the first
    # num_sequences_to subjects (functions.cpp:125-149)
    first, count = 100, 250
    orc = oracle_run(q, res[offs[first]:offs[first + count]], offs[first:first + count + 1] - offs[first],
                     lens[first:first + count])
    assert_parity(gpu_run(ctx, q, res, offs, lens, first=first, count=count), orc, first)
    # nothing to scan
    st, recs, subs, ustats, _ = gpu_run(ctx, q, res, offs, lens, first=10, count=0)
    assert len(st) == 0 and len(recs) == 0 and ustats["lookup_hits"] == 0


def test_overlapping_scans_keep_their_own_counters(ctx):
    """The kernel writes per-subject counters straight to pinned host memory.
    Two scans in flight at once (the context's two stream lanes) each fill
    their own buffer; a third scan, started while the second is pending,
    reuses the buffer the first released."""
    rng = np.random.default_rng(21)
    qa, qb = rng.choice(STD20, 300), rng.choice(STD20, 180)
    subjects = [np.concatenate([rng.choice(STD20, int(rng.integers(0, 400))),
                                (qa if i % 2 else qb)[: int(rng.integers(20, 180))]]) for i in range(3000)]
    res, offs, lens = pack(subjects)
    say("overlapping scans: 3,000 subjects, queries of 300 and 180")
    db = ctx.load_db(res, offs[:-1], lens)
    da, dq_b = ctx.prepare_query(qa), ctx.prepare_query(qb)

    def grab(r):
        r.wait()
        recs, subs = r.records()
        st = r.subject_stats()
        _, u = r.summary()
        r.close()
        return st, recs, subs, u, None

    ra = bpx.scan(ctx, da, db, subject_stats=True)
    rb = bpx.scan(ctx, dq_b, db, subject_stats=True)
    ga = grab(ra)
    rc = bpx.scan(ctx, da, db, subject_stats=True)
    gb = grab(rb)
    gc = grab(rc)
    orc_a = oracle_run(qa, res, offs, lens)
    assert_parity(ga, orc_a)
    assert_parity(gb, oracle_run(qb, res, offs, lens))
    assert_parity(gc, orc_a)
    da.close()
    dq_b.close()
    db.close()


def test_long_subjects(ctx):
    """Subjects longer than 65,536 residues exercise the 16-bit ring rebase."""
    rng = np.random.default_rng(6)
    q = rng.choice(STD20, 400)
    subjects = []
    for n in (35000, 70000, 140001):
        s = rng.choice(STD20, n)
        for at in range(1000, n - 500, 9000):
            s[at:at + 400] = q
        subjects.append(s)
    res, offs, lens = pack(subjects)
    assert_parity(gpu_run(ctx, q, res, offs, lens), oracle_run(q, res, offs, lens))


@pytest.mark.parametrize("qlen", [3, 50, 500, 2000, 4000])
def test_query_lengths_and_table_placement(ctx, qlen):
    rng = np.random.default_rng(qlen)
    q = bpx.synth_query(11, qlen, qlen)
    subjects = [rng.choice(STD20, int(rng.integers(20, 900))) for _ in range(1500)]
    for i in range(0, 1500, 50):
        subjects[i] = np.concatenate([subjects[i], q[: min(qlen, 300)]])
    res, offs, lens = pack(subjects)
    gpu = gpu_run(ctx, q, res, offs, lens)
    say(f"qlen {qlen}: table {gpu[4]}")
    assert_parity(gpu, oracle_run(q, res, offs, lens))


def test_word_length_2_and_custom_matrix(ctx):
    rng = np.random.default_rng(8)
    M = rng.integers(-4, 7, (28, 28)).astype(np.int8)
    M[0, :] = -128
    M[:, 0] = -128
    q = rng.choice(STD20, 90)
    subjects = [np.concatenate([rng.choice(STD20, int(rng.integers(0, 200))), q[::-1], q]) for _ in range(400)]
    res, offs, lens = pack(subjects)
    for word, T in ((2, 8), (3, 10)):
        p = bpx.default_params(word_length=word, threshold=T, cutoff_score=30)
        assert_parity(gpu_run(ctx, q, res, offs, lens, p, M), oracle_run(q, res, offs, lens, p, M))


def test_segment_overflow_rescan(ctx):
    """Every subject carries many high-scoring copies of the query: the
    per-wave record segments overflow and the library must rescan with
    larger ones, returning exactly the oracle's records."""
    rng = np.random.default_rng(9)
    q = rng.choice(STD20, 60)
    unit = np.concatenate([q, rng.choice(STD20, 45)])
    subjects = [np.tile(unit, int(rng.integers(60, 120))) for _ in range(6000)]
    res, offs, lens = pack(subjects)
    # a fresh context (no segment size learned yet) and one wave per CU, so
    # each wave collects far more records than the first-scan segment holds
    fresh = bpx.Context(0)
    gpu = gpu_run(fresh, q, res, offs, lens, launch=dict(waves_per_group=1))
    fresh.close()
    say(f"overflow case: {len(gpu[1])} records")
    assert len(gpu[1]) > 6000 * 60 // 2
    assert_parity(gpu, oracle_run(q, res, offs, lens))


@pytest.mark.parametrize("launch", [dict(waves_per_group=1), dict(waves_per_group=4, groups_per_cu=2),
                                    dict(waves_per_group=8, subjects_per_item=8, subject_buffer=1024),
                                    dict(queue_capacity=4096), dict(subject_buffer=1024),
                                    dict(subject_buffer=8192, waves_per_group=4)])
def test_launch_shapes_do_not_change_results(ctx, launch):
    rng = np.random.default_rng(10)
    q = bpx.synth_query(3, 0, 350)
    subjects = [rng.choice(STD20, int(rng.integers(30, 1200))) for _ in range(3000)]
    for i in range(0, 3000, 30):
        subjects[i][: min(len(subjects[i]), 200)] = q[: min(len(subjects[i]), 200)]
    res, offs, lens = pack(subjects)
    assert_parity(gpu_run(ctx, q, res, offs, lens, launch=launch), oracle_run(q, res, offs, lens))


@pytest.mark.parametrize("launch", [None, {"waves_per_group": 1}, {"subject_buffer": 1024}])
def test_low_complexity_chains_and_queue_caps(ctx, launch):
    """Homopolymer runs and tandem repeats against a repetitive query: long
    chains of hits on one diagonal (more than two claim rounds, so hits ride
    over to the next batch), cells larger than a batch, and more pairs and
    deferred hits than the per-wave queues hold (flushes in mid-subject)."""
    rng = np.random.default_rng(91)
    motif = rng.choice(STD20, 7)
    q = np.concatenate([np.tile(motif, 30), np.full(60, STD20[3], np.uint8), rng.choice(STD20, 90)])
    subjects = []
    for i in range(400):
        kind = i % 4
        n = int(rng.integers(40, 2600))
        if kind == 0:
            s = np.full(n, STD20[3], np.uint8)                       # homopolymer
        elif kind == 1:
            s = np.tile(motif, n // 7 + 1)[:n].copy()                # tandem repeat of the query's motif
        elif kind == 2:
            s = np.tile(rng.choice(STD20, int(rng.integers(2, 6))), n // 2 + 1)[:n].copy()
        else:
            s = rng.choice(STD20, n)
            at = int(rng.integers(0, max(1, n - 200)))
            s[at:at + 200] = np.tile(motif, 29)[: len(s[at:at + 200])]
        mut = rng.random(n) < 0.05
        s[mut] = rng.choice(STD20, int(mut.sum()))
        subjects.append(s)
    res, offs, lens = pack(subjects)
    say(f"low complexity, launch {launch}: {int(lens.sum())} residues")
    gpu = gpu_run(ctx, q, res, offs, lens, launch=launch)
    assert_parity(gpu, oracle_run(q, res, offs, lens))
    # runs of more than 64 records per subject take the compaction's heapsort
    assert int(np.bincount(gpu[2]).max()) > 64


def test_infeasible_launch_is_refused(ctx):
    q = bpx.synth_query(3, 1, 350)
    res, offs, lens = pack([q, q])
    db = ctx.load_db(res, offs[:-1], lens)
    dq = ctx.prepare_query(q)
    with pytest.raises(bpx.BpxError, match="LDS plan"):
        bpx.scan(ctx, dq, db, launch=dict(waves_per_group=8, subject_buffer=65536))
    rc, dev, lds = bpx.check_memory(ctx, dq, db, launch=dict(waves_per_group=8, subject_buffer=65536))
    assert rc == -5 and lds > ctx.info()["lds_bytes"]


def test_device_memory_limit_refuses_a_scan_that_would_not_fit(ctx):
    """GPU_BLASTP_check_memory (gpu_blastp.cu:316-333) compares the sum of a
    search's buffers with the device size; with the context limited below
    that sum, bpx_check_memory must return BPX_E_BUDGET, and pass again once
    the limit is lifted."""
    q = bpx.synth_query(4, 1, 300)
    spec = dict(seed=8, num_subjects=2000, len_median=300, len_sigma=0.6, len_min=30, len_max=3000)
    sdb = bpx.synth_db(spec, [q], nthreads=THREADS)
    db = ctx.load_db(sdb.residues, sdb.offsets[:-1], sdb.lengths)
    dq = ctx.prepare_query(q)
    rc, need, _ = bpx.check_memory(ctx, dq, db)
    assert rc == 0 and need > 0
    try:
        ctx.limit(need - 1)
        rc, need2, _ = bpx.check_memory(ctx, dq, db)
        assert rc == -5 and need2 == need
        assert "device bytes" in bpx.last_error()
    finally:
        ctx.limit(0)
    assert bpx.check_memory(ctx, dq, db)[0] == 0
    dq.close()
    db.close()


def test_refused_pool_allocation_is_retried_without_a_stale_error(monkeypatch):
    """The device pool retries an allocation HIP refuses after releasing its
    cached blocks.  The refusal is handled there, so it must not reach the
    next kernel launch check, which reads the thread's last HIP error (a
    stale out-of-memory once failed an autotuner trial that way).  With
    BPX_POOL_FAIL_FIRST set, every device-pool allocation of a new context
    first asks HIP for 4 EiB: a plain and a batched scan run through that
    path and still equal the oracle, also after bpx_context_trim has freed
    the cache."""
    monkeypatch.setenv("BPX_POOL_FAIL_FIRST", "100000")
    c = bpx.Context(0)
    try:
        q = bpx.synth_query(6, 1, 250)
        spec = dict(seed=9, num_subjects=1500, len_median=300, len_sigma=0.6, len_min=30, len_max=3000)
        sdb = bpx.synth_db(spec, [q], nthreads=THREADS)
        res, offs, lens = sdb.residues, sdb.offsets, sdb.lengths
        orc = oracle_run(q, res, offs, lens)
        assert_parity(gpu_run(c, q, res, offs, lens), orc)
        c.trim()  # the cache freed: the next scan allocates afresh
        assert_parity(gpu_run(c, q, res, offs, lens), orc)
        queries = [bpx.synth_query(26, i, L) for i, L in enumerate([40, 120, 300, 700, 1500])]
        assert_batch_parity(c, queries, batch_db(queries, n=800, seed=36), group_residues=1000, group_queries=4)
    finally:
        c.close()


def test_query_footprint(ctx):
    """The footprint of a query scanned alone: its diagonal slots (next power
    of two >= qlen + window), the waves it runs at with its table in HBM; a
    query needing more than 65,536 slots is refused."""
    spec = dict(seed=9, num_subjects=300, len_median=300, len_sigma=0.6, len_min=30, len_max=3000)
    sdb = bpx.synth_db(spec)
    db = ctx.load_db(sdb.residues, sdb.offsets[:-1], sdb.lengths)
    rc, slots, waves, lds = bpx.query_footprint(ctx, db, 500)
    assert rc == 0 and slots == 1024 and waves == 16 and 0 < lds <= ctx.info()["lds_bytes"]
    rc, slots, waves, _ = bpx.query_footprint(ctx, db, 30000)
    assert rc == 0 and slots == 32768 and 1 <= waves < 16
    rc, slots, waves, _ = bpx.query_footprint(ctx, db, 65500)
    assert rc == -5 and slots == 131072 and waves == 0
    db.close()


def test_database_load_checks_codes_and_handles_empty(ctx):
    """The arena is repacked by host threads in chunks; the first bad residue
    code (lowest subject, then position) is reported, and an empty volume
    loads and scans to nothing."""
    rng = np.random.default_rng(22)
    subjects = [rng.choice(STD20, int(rng.integers(1, 400))) for _ in range(20_000)]
    subjects[9_000][3] = 28
    subjects[15_000][0] = 200
    subjects[9_000][7] = 31
    res, offs, lens = pack(subjects)
    with pytest.raises(bpx.BpxError, match="subject 9000 position 3: residue code 28"):
        ctx.load_db(res, offs[:-1], lens)
    empty = ctx.load_db(np.zeros(0, np.uint8), np.zeros(0, np.uint64), np.zeros(0, np.uint32))
    q = ctx.prepare_query(rng.choice(STD20, 100))
    r = bpx.scan(ctx, q, empty, subject_stats=True).wait()
    n, ust = r.summary()
    assert n == 0 and len(r.subject_stats()) == 0 and ust["lookup_hits"] == 0
    r.close()
    q.close()
    empty.close()


def workload_db(which):
    if which == 1:
        q = bpx.synth_query(0, 0, 300)
        spec = dict(seed=0, num_subjects=10_000, len_median=300, len_sigma=0.6, len_min=30, len_max=5000,
                    homolog_count=50, identity_min=0.5, identity_max=0.5, indel_min=0.02, indel_max=0.02)
    else:
        q = bpx.synth_query(0, 0, 500)
        spec = dict(seed=2, num_subjects=2_000_000, len_median=300, len_sigma=0.6, len_min=30, len_max=5000,
                    homolog_fraction=0.005, identity_min=0.3, identity_max=0.7, indel_min=0.0, indel_max=0.03)
    return q, bpx.synth_db(spec, [q], nthreads=THREADS)


def test_10k_subject_workload_full(ctx):
    q, db = workload_db(1)
    say(f"10,000-subject workload: {db.num_subjects} subjects, {db.num_residues} residues, sha256 {db.sha256}")
    gpu = gpu_run(ctx, q, db.residues, db.offsets, db.lengths)
    orc = oracle_run(q, db.residues, db.offsets, db.lengths)
    assert_parity(gpu, orc)
    assert len(orc[2]) > 0


def test_2M_subject_workload_full(ctx):
    q, db = workload_db(2)
    say(f"2,000,000-subject workload: {db.num_subjects} subjects, {db.num_residues} residues, sha256 {db.sha256}")
    gpu = gpu_run(ctx, q, db.residues, db.offsets, db.lengths)
    say(f"2,000,000-subject workload: GPU done, {len(gpu[1])} records")
    orc = oracle_run(q, db.residues, db.offsets, db.lengths)
    say(f"2,000,000-subject workload: oracle done, {len(orc[2])} records")
    assert_parity(gpu, orc)


def batch_db(queries, n=3000, seed=31):
    spec = dict(seed=seed, num_subjects=n, len_median=300, len_sigma=0.8, len_min=20, len_max=6000,
                homolog_fraction=0.05, identity_min=0.4, identity_max=0.7, indel_min=0.0, indel_max=0.03,
                biased_fraction=0.1, dirichlet_alpha=0.3, motif_min=5, motif_max=30)
    return bpx.synth_db(spec, [q for q in queries if len(q) > 0], nthreads=THREADS)


def assert_batch_parity(ctx, queries, db, xdrop=None, cutoff=None, launch=None, **policy):
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    b = ctx.prepare_batch(queries, x_dropoff=xdrop, cutoff=cutoff, **policy)
    say(f"batch: {b.info()}")
    r = bpx.scan_batch(ctx, b, ddb, subject_stats=True, launch=launch).wait()
    recs, subs, qids = r.records(with_queries=True)
    st = r.subject_stats().reshape(len(queries), db.num_subjects)
    qstats = r.query_stats()
    assert (np.diff(qids.astype(np.int64)) >= 0).all()
    for i, q in enumerate(queries):
        prm = O.params(len(q), x_dropoff=int(xdrop[i]) if xdrop is not None else 16,
                       cutoff=int(cutoff[i]) if cutoff is not None else 41)
        orc = O.run_db(q, db.residues, db.offsets, db.lengths, prm, THREADS)
        sel = qids == i
        gpu = (st[i], recs[sel], subs[sel], qstats[i], None)
        assert_parity(gpu, orc)
    r.close()
    b.close()
    ddb.close()


def test_batch_equals_single_queries(ctx):
    rng = np.random.default_rng(21)
    lens = [int(x) for x in rng.integers(5, 900, 40)] + [2000, 3, 0, 1]
    queries = [bpx.synth_query(21, i, L) for i, L in enumerate(lens)]
    db = batch_db(queries)
    assert_batch_parity(ctx, queries, db)


def test_batch_per_query_parameters_and_policies(ctx):
    rng = np.random.default_rng(22)
    lens = [int(x) for x in rng.integers(20, 400, 30)]
    queries = [bpx.synth_query(22, i, L) for i, L in enumerate(lens)]
    db = batch_db(queries, n=1500, seed=32)
    xd = rng.integers(8, 30, len(queries)).astype(np.int32)
    co = rng.integers(25, 50, len(queries)).astype(np.int32)
    for pol in (dict(group_residues=300, group_queries=4), dict(group_residues=4000, group_queries=200),
                dict(group_residues=4000, group_queries=64, group_diag_slots=1536), dict(group_residues=1)):
        say(f"policy {pol}")
        assert_batch_parity(ctx, queries, db, xd, co, **pol)


@pytest.mark.parametrize("launch", [dict(claims_per_share=1), dict(waves_per_group=1, claims_per_share=2),
                                    dict(subjects_per_item=8)])
def test_batch_multi_subject_claims_match_oracle(ctx, launch):
    """Packed groups on a database large enough that a wavefront claims
    several subjects at a time (claims_per_share 1 and 2 give chunks of up to
    16), and with explicit 8-subject work items: every query equals the
    oracle."""
    rng = np.random.default_rng(25)
    lens = [int(x) for x in rng.integers(30, 700, 24)]
    queries = [bpx.synth_query(25, i, L) for i, L in enumerate(lens)]
    db = batch_db(queries, n=30000, seed=35)
    say(f"launch {launch}")
    assert_batch_parity(ctx, queries, db, launch=launch, group_residues=2000, group_queries=8)


def test_autotuned_batch_matches_oracle(ctx):
    """The autotuner only reads timings: whatever it picks, the batched scan
    under that packing and launch shape equals the oracle per query."""
    rng = np.random.default_rng(23)
    lens = [int(x) for x in np.exp(rng.uniform(np.log(50), np.log(2000), 24))]
    queries = [bpx.synth_query(23, i, L) for i, L in enumerate(lens)]
    db = batch_db(queries, n=2000, seed=33)
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    policy, launch, trials = bpx.autotune_batch(ctx, queries, ddb, 100, 400)
    say(f"autotune: {policy} {launch}; {len(trials)} trials")
    assert len(trials) >= 4 and all(t["sample_ms"] > 0 for t in trials)
    ddb.close()
    assert_batch_parity(ctx, queries, db, launch=launch, **policy)


def test_graph_replay_equals_eager(ctx):
    """A graph-mode batched scan replayed several times, with eager scans in
    between, gives the eager result on every replay (records, subject and
    query ids, counters, per-query stats)."""
    rng = np.random.default_rng(24)
    lens = [int(x) for x in rng.integers(20, 700, 30)]
    queries = [bpx.synth_query(24, i, L) for i, L in enumerate(lens)]
    db = batch_db(queries, n=2500, seed=34)
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    b = ctx.prepare_batch(queries, group_residues=900, group_queries=16)

    def snapshot(r):
        recs, subs, qids = r.records(with_queries=True)
        return (recs.view(np.int32).copy(), subs.copy(), qids.copy(), r.subject_stats().view(np.int32).copy(),
                [tuple(d.values()) for d in r.query_stats()])

    eager = bpx.scan_batch(ctx, b, ddb, subject_stats=True).wait()
    want = snapshot(eager)
    eager.close()
    g = bpx.scan_batch(ctx, b, ddb, subject_stats=True, graph=True).wait()
    for rep in range(4):
        got = snapshot(g)
        for x, y in zip(got[:4], want[:4]):
            np.testing.assert_array_equal(x, y)
        assert got[4] == want[4], rep
        # eager work and a synchronisation between replays
        other = bpx.scan_batch(ctx, b, ddb, first=100, count=500).wait()
        other.close()
        g.replay().wait()
    g.close()
    b.close()
    ddb.close()


def test_volumes_load_as_one_arena(ctx, tmp_path):
    """An alias of three NCBI volumes loads as one arena (subject i = OID i)
    and scans exactly like the same subjects given as arrays."""
    q = bpx.synth_query(60, 0, 420)
    spec = dict(seed=61, num_subjects=3000, len_median=300, len_sigma=0.6, len_min=30, len_max=5000,
                homolog_fraction=0.03, identity_min=0.4, identity_max=0.7, indel_min=0.0, indel_max=0.02)
    db = bpx.synth_db(spec, [q], nthreads=THREADS)
    cuts = [0, 1000, 1001, 3000]
    names = []
    for i in range(3):
        a, b = cuts[i], cuts[i + 1]
        o = db.offsets[a:b + 1] - db.offsets[a]
        names.append(f"syn.{i:02d}")
        bpx.volume_write(str(tmp_path / names[-1]), db.residues[db.offsets[a]:db.offsets[b]], o[:-1],
                         db.lengths[a:b])
    (tmp_path / "syn.pal").write_text("TITLE synthetic\nDBLIST " + " ".join(names) + "\n")
    vdb = bpx.load_volumes(ctx, str(tmp_path / "syn"))
    assert vdb.num_subjects == db.num_subjects and vdb.num_residues == db.num_residues
    dq = ctx.prepare_query(q)
    r = bpx.scan(ctx, dq, vdb).wait()
    recs, subs = r.records()
    st = r.subject_stats()
    _, ustats = r.summary()
    r.close()
    assert_parity((st, recs, subs, ustats, None), oracle_run(q, db.residues, db.offsets, db.lengths))
    dq.close()
    vdb.close()


def test_imported_reference_gpu_volume_scans_like_the_oracle(ctx, tmp_path):
    """A database in the reference's own .gpu / .gpuinfo files -- laid out by
    the test suite's restatement of its writer (functions.cpp:364-613, 1-46)
    -- is imported (bpx_gpu_volume_read), loaded and scanned on the GPU, and
    equals the oracle on the same sequences."""
    from test_reference_gpu_db import write_reference_files
    q = bpx.synth_query(62, 0, 380)
    spec = dict(seed=63, num_subjects=3000, len_median=300, len_sigma=0.6, len_min=30, len_max=4000,
                homolog_fraction=0.03, identity_min=0.4, identity_max=0.7, indel_min=0.0, indel_max=0.02)
    sdb = bpx.synth_db(spec, [q], nthreads=THREADS)
    order = np.argsort(sdb.lengths, kind="stable")  # the reference's files hold sequences by ascending length
    seqs = [sdb.residues[int(sdb.offsets[i]):int(sdb.offsets[i + 1])] for i in order]
    path = write_reference_files(str(tmp_path), "ref", [(seqs, len(seqs))], blocks=8, threads=32)
    res, offs, lens, _ = bpx.gpu_volume_read(path, 0)
    assert len(lens) == len(seqs)
    db = ctx.load_db(res, offs[:-1], lens)
    dq = ctx.prepare_query(q)
    r = bpx.scan(ctx, dq, db).wait()
    recs, subs = r.records()
    st = r.subject_stats()
    _, ustats = r.summary()
    r.close()
    ores, ooffs, olens = pack(seqs)
    assert_parity((st, recs, subs, ustats, None), oracle_run(q, ores, ooffs, olens))
    assert len(recs) > 0
    dq.close()
    db.close()


def test_phase_times_are_consistent_and_change_nothing(ctx):
    """BPX_OUT_PHASES: every subject is staged once per group pass, at least
    its residues' bytes; wavefront lifetimes lie within the launch span and
    contain the staging and group-change times; the results are those of the
    same scan without phase times."""
    q, db = workload_db(1)
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    dq = ctx.prepare_query(q)
    plain = bpx.scan(ctx, dq, ddb).wait()
    want = plain.records()[0].view(np.int32).copy()
    plain.close()
    r = bpx.scan(ctx, dq, ddb, phases=True).wait()
    np.testing.assert_array_equal(r.records()[0].view(np.int32), want)
    ph = r.phases()
    say(f"phases: {ph}")
    nw = ctx.info()["num_cus"] * ph["waves"] // ctx.info()["num_cus"]
    assert ph["waves"] == nw and ph["waves"] % ctx.info()["num_cus"] == 0
    assert ph["subjects_staged"] == db.num_subjects
    assert db.num_residues <= ph["bytes_staged"] <= db.num_residues + 16 * db.num_subjects
    assert 0 < ph["stage_ms_mean"] <= ph["wave_ms_mean"] <= ph["span_ms"] * 1.001
    assert 0 <= ph["group_ms_mean"] <= ph["wave_ms_mean"]
    # every extension the scan ran went through a flush
    assert ph["extensions"] == r.summary()[1]["init_extends"] > 0
    assert 0 < ph["flush_ms_mean"] < ph["wave_ms_mean"]
    # one launch class: the busy fraction is the mean lifetime over the span
    assert ph["classes"] == 1
    assert abs(ph["busy_fraction"] - ph["wave_ms_mean"] / ph["span_ms"]) < 1e-6
    assert 0 < ph["busy_fraction"] <= 1.0 + 1e-9
    r.close()
    dq.close()
    ddb.close()

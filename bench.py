"""Benchmark of the gfx950 blastp word finder (two-hit word finding +
ungapped X-drop extension) on one MI355X.

--workload plain (default): one step = one pass of the hot path over one
batch: the 500-residue query scanned against the resident 2,000,000-subject
database (~720 M residues, lognormal lengths median 300 sigma 0.6, 0.5 %
planted homologs at 30-70 % identity, 0-3 % indels), including the record
compaction and the copy of records and per-subject counters back to the host.

--workload batched: 1,024 queries (log-uniform 50-2,000 aa) against a
500,000-subject database (1 % homologs of random queries at 40-60 %
identity); one step = one batched scan of all queries (autotuned packing and
launch shape), compared with scanning the queries one by one.

--workload skewed: the same 1,024 queries against a skewed 500,000-subject
database (lognormal lengths median 350 sigma 1.2, clipped 20-35,000; 20 %
compositionally biased with tandem repeats); the learned cost model
(bpx.costsched) orders and bins the work; one step = one batched scan,
compared with the static longest-first order.

--workload serving: single queries (lognormal lengths, median 350, sigma
0.6, clipped 20-3,000) arriving as a Poisson stream against the resident
2,000,000-subject database, served by the continuous-batching server
(bpx.Server); reports throughput and p50/p99 latency per arrival rate, and
checks every result against the plain single-query scan.

Inputs are resident in HBM before the timed region (database arena and query
tables uploaded once).  Prints ONE JSON line; extra
detail goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "blastp-cuda_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "DB residues/s on 1 GPU, batched+binned vs plain; HBM GB/s; p50/p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def plain_workload(threads):
    import bpx
    q = bpx.synth_query(0, 0, 500)
    spec = dict(seed=2, num_subjects=2_000_000, len_median=300, len_sigma=0.6, len_min=30, len_max=5000,
                homolog_fraction=0.005, identity_min=0.3, identity_max=0.7, indel_min=0.0, indel_max=0.03)
    return q, bpx.synth_db(spec, [q], nthreads=threads)


def batched_workload(threads):
    import bpx
    lens = bpx.synth_query_lengths(3, 1024, "loguniform", 50, 2000)
    qs = [bpx.synth_query(3, i, int(n)) for i, n in enumerate(lens)]
    spec = dict(seed=3, num_subjects=500_000, len_median=300, len_sigma=0.6, len_min=30, len_max=5000,
                homolog_fraction=0.01, identity_min=0.4, identity_max=0.6, indel_min=0.0, indel_max=0.03)
    return qs, bpx.synth_db(spec, qs, nthreads=threads)


def cpu_baseline(q, db, threads, budget_s=20.0):
    """The C oracle (same semantics, same inputs) on bounded prefixes of the
    database, timed on this host's cores: on `threads` threads, the median of
    5 runs after a warm-up; on one thread, the median of 3 runs on a smaller
    prefix after a warm-up."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    lk = O.OracleLookup(q)

    def timed(nthreads, share, runs):
        # size the prefix from a short probe so this leg takes ~share of budget_s
        probe = 2000
        t = time.perf_counter()
        O.run_db(q, db.residues, db.offsets[: probe + 1], db.lengths[:probe], nthreads=nthreads, lookup=lk)
        dt = max(time.perf_counter() - t, 1e-4)
        rate = db.offsets[probe] / dt
        target = int(rate * budget_s * share / (runs + 1))
        n = int(np.searchsorted(db.offsets, min(target, int(db.offsets[-1]))))
        n = max(probe, min(n, db.num_subjects))
        times = []
        for _ in range(runs + 1):
            t = time.perf_counter()
            O.run_db(q, db.residues, db.offsets[: n + 1], db.lengths[:n], nthreads=nthreads, lookup=lk)
            times.append(time.perf_counter() - t)
        res = int(db.offsets[n])
        return n, res, res / float(np.median(times[1:]))

    n, res, value = timed(threads, 0.6, 5)
    n1, res1, value1 = timed(1, 0.4, 3)
    lk.close()
    return {"value": value, "unit": "DB residues/s", "cores": threads, "kind": "port",
            "sample": f"first {n} subjects ({res} residues) of the same database and query; "
                      f"median of 5 runs after a warm-up, {threads} threads",
            "value_1thread": value1,
            "sample_1thread": f"first {n1} subjects ({res1} residues); median of 3 runs after a warm-up, 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 20; batched 5)")
    ap.add_argument("--warmup", type=int, default=None, help="warm-up steps (default 3; batched 1)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--workload", choices=("plain", "batched", "skewed", "serving"), default="plain")
    ap.add_argument("--rates", default="25,50,200,2000", help="serving: arrival rates (queries/s), comma-separated")
    ap.add_argument("--requests", type=int, default=120, help="serving: requests per arrival rate")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="first run a plain elementwise copy (torch.neg over int32) reading and writing BYTES, "
                         "3 times: the known-size kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"batched": 5, "skewed": 2}.get(args.workload, 20)
    if args.warmup is None:
        args.warmup = 1 if args.workload in ("batched", "skewed") else 3
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the word finder shards by "
              "subject across GPUs, but this build measures the single-GPU path.", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe > 0:
        import torch
        n = args.copy_probe // 4
        a = torch.arange(n, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.neg(a, out=b)
        torch.cuda.synchronize()
        log(f"copy probe: {3} x torch.neg over {n} int32 ({4 * n} B read, {4 * n} B written per launch)")
        del a, b

    if args.workload == "batched":
        return batched(args)
    if args.workload == "skewed":
        return skewed(args)
    if args.workload == "serving":
        return serving(args)

    import bpx
    log("generating the 500-aa x 2,000,000-subject workload ...")
    t0 = time.perf_counter()
    q, db = plain_workload(args.threads)
    log(f"  {db.num_subjects} subjects, {db.num_residues} residues, sha256 {db.sha256} "
        f"({time.perf_counter() - t0:.1f} s)")
    ctx = bpx.Context(0)
    dinfo = ctx.info()
    t_load = time.perf_counter()
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    load_s = time.perf_counter() - t_load
    dq = ctx.prepare_query(q)
    log(f"  device {dinfo}; query table {dq.info()}; database load (host arena + upload) {load_s:.3f} s")

    def step():
        r = bpx.scan(ctx, dq, ddb, subject_stats=True).wait()
        t = r.timing()
        n, st = r.summary()
        r.close()
        return t, n, st

    for _ in range(args.warmup):
        step()
    kernel_ms, device_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t, nrec, ust = step()
        kernel_ms.append(t["kernel_ms"])
        device_ms.append(t["device_ms"])
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps
    value = db.num_residues / (ms_per_step / 1000.0)

    # roofline of the word-finder kernel: algorithmic HBM bytes per launch =
    # residues (1 B) + subject descriptors (12 B) read + records (24 B)
    # written; the per-subject counters (12 B) go to pinned host memory
    kern = float(np.mean(kernel_ms))
    bytes_alg = db.num_residues + 12 * db.num_subjects + 24 * nrec
    achieved = bytes_alg / (kern / 1000.0) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "k_scan<3,true,false,true>", "kernel_ms_mean": round(kern, 4),
                "kernel_ms_p50": round(float(np.percentile(kernel_ms, 50)), 4),
                "kernel_ms_p99": round(float(np.percentile(kernel_ms, 99)), 4),
                "bytes_alg": int(bytes_alg)}
    traffic_file = os.path.join(ROOT, "profiles", "traffic_k_scan_plain_500aa_2M_subjects.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            roofline["traffic"] = json.load(f).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU oracle on a bounded sample ...")
        cpu = cpu_baseline(q, db, args.threads)

    out = {"metric": METRIC, "value": round(value, 1), "unit": "DB residues/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int8/int32",
           "data": "synthetic (seeded generator, BLOSUM62 background residues, planted homologs)",
           "config": {"workload": "plain: 1 query x 500 aa vs 2,000,000-subject DB",
                      "queries": 1, "query_len": 500, "subjects": db.num_subjects,
                      "residues": db.num_residues, "db_sha256": db.sha256,
                      "records": int(nrec), "lookup_hits": ust["lookup_hits"],
                      "init_extends": ust["init_extends"], "device_ms_mean": round(float(np.mean(device_ms)), 4),
                      # a search that first loads the volume from host memory (not `value`)
                      "db_load_s": round(load_s, 4),
                      "cold_search_residues_per_s": round(db.num_residues / (load_s + ms_per_step / 1000.0), 1)},
           "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def cpu_baseline_batched(qs, db, threads, budget_s=20.0):
    """The C oracle over every 16th query of the batch against a bounded
    prefix of the database, scaled to query-residue pairs per second."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    sample_q = qs[::16]
    probe = 500
    t = time.perf_counter()
    for q in sample_q:
        O.run_db(q, db.residues, db.offsets[: probe + 1], db.lengths[:probe], nthreads=threads)
    dt = max(time.perf_counter() - t, 1e-4)
    n = int(min(db.num_subjects, max(probe, probe * budget_s / 2 / dt)))
    t = time.perf_counter()
    for q in sample_q:
        O.run_db(q, db.residues, db.offsets[: n + 1], db.lengths[:n], nthreads=threads)
    dt = time.perf_counter() - t
    res = int(db.offsets[n]) * len(sample_q)
    return {"value": res / dt, "unit": "DB residues/s", "cores": threads, "kind": "port",
            "sample": f"{len(sample_q)} of the 1,024 queries (every 16th) against the first {n} subjects; "
                      f"one timed run after a probe, {threads} threads"}


def batched(args):
    import bpx
    log("generating the 1,024-query batched workload ...")
    t0 = time.perf_counter()
    qs, db = batched_workload(args.threads)
    qres = int(sum(len(q) for q in qs))
    log(f"  {len(qs)} queries ({qres} residues), {db.num_subjects} subjects, {db.num_residues} residues, "
        f"sha256 {db.sha256} ({time.perf_counter() - t0:.1f} s)")
    ctx = bpx.Context(0)
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    work = db.num_residues * len(qs)  # DB residues summed over queries

    # autotune on a 10 % sample (subjects [0, N/10))
    t = time.perf_counter()
    gr, gq, launch, trials = bpx.autotune_batch(ctx, qs, ddb, 0, db.num_subjects // 10)
    log(f"  autotune: group_residues={gr} group_queries={gq} launch={launch} ({time.perf_counter() - t:.1f} s, "
        f"{len(trials)} trials)")
    for tr in trials:
        log(f"    {tr}")
    b = ctx.prepare_batch(qs, group_residues=gr, group_queries=gq)
    info = b.info()
    # one HIP graph holds the whole batched scan (counter resets, every launch
    # class, compaction, result copies); each step replays it
    graph = bpx.scan_batch(ctx, b, ddb, launch=launch, graph=True).wait()

    def step():
        graph.replay().wait()
        return graph.timing(), graph.summary()[0]

    for _ in range(args.warmup):
        step()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tm, nrec = step()
        kernel_ms.append(tm["kernel_ms"])
    ms_per_step = (time.perf_counter() - t0) * 1000.0 / args.steps
    value = work / (ms_per_step / 1000.0)
    graph.close()

    # plain: every query on its own (one warm-up pass over 32 queries)
    dqs = [ctx.prepare_query(q) for q in qs]
    for dq in dqs[:32]:
        bpx.scan(ctx, dq, ddb, subject_stats=False).wait().close()
    t0 = time.perf_counter()
    plain_records = 0
    for dq in dqs:
        r = bpx.scan(ctx, dq, ddb, subject_stats=False).wait()
        plain_records += r.summary()[0]
        r.close()
    plain_s = time.perf_counter() - t0
    plain_value = work / plain_s
    if plain_records != nrec:
        raise SystemExit(f"batched and plain record counts differ: {nrec} vs {plain_records}")
    log(f"  batched {ms_per_step:.1f} ms/step -> {value / 1e9:.2f} G res/s; plain {plain_s:.2f} s -> "
        f"{plain_value / 1e9:.2f} G res/s; ratio {value / plain_value:.2f}")

    kern = float(np.mean(kernel_ms))
    passes = info["groups"]
    # every group pass reads the residues and descriptors; each (query,
    # subject) pair's record count is zero-filled, summed and turned into its
    # CSR offset in place (12 B); records are 24 B
    bytes_alg = passes * (db.num_residues + 12 * db.num_subjects) + 12 * db.num_subjects * len(qs) + 24 * nrec
    achieved = bytes_alg / (kern / 1000.0) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "k_scan (the graph of the batched scan: all launch classes + compaction)",
                "kernel_ms_mean": round(kern, 3),
                "bytes_alg": int(bytes_alg), "passes": passes}
    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU oracle on a bounded sample ...")
        cpu = cpu_baseline_batched(qs, db, args.threads)
    out = {"metric": METRIC, "value": round(value, 1), "unit": "DB residues/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int8/int32",
           "data": "synthetic (seeded generator, BLOSUM62 background residues, planted homologs)",
           "config": {"workload": "batched: 1,024 queries (log-uniform 50-2,000 aa) vs 500,000-subject DB",
                      "queries": len(qs), "query_residues": qres, "subjects": db.num_subjects,
                      "residues": db.num_residues, "db_sha256": db.sha256, "records": int(nrec),
                      "groups": passes, "group_residues": gr, "group_queries": gq, "launch": launch,
                      "plain_value": round(plain_value, 1), "plain_seconds": round(plain_s, 3),
                      "batched_over_plain": round(value / plain_value, 3)},
           "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


SKEWED_SPEC = dict(seed=4, num_subjects=500_000, len_median=350, len_sigma=1.2, len_min=20, len_max=35000,
                   homolog_fraction=0.01, identity_min=0.4, identity_max=0.6, indel_min=0.0, indel_max=0.03,
                   biased_fraction=0.2, dirichlet_alpha=0.3, motif_min=5, motif_max=30)


def skewed(args):
    import bpx
    from bpx import costsched
    log("generating the skewed-database workload ...")
    t0 = time.perf_counter()
    lens = bpx.synth_query_lengths(3, 1024, "loguniform", 50, 2000)
    qs = [bpx.synth_query(3, i, int(n)) for i, n in enumerate(lens)]
    db = bpx.synth_db(SKEWED_SPEC, qs, nthreads=args.threads)
    log(f"  {len(qs)} queries, {db.num_subjects} subjects, {db.num_residues} residues, sha256 {db.sha256} "
        f"({time.perf_counter() - t0:.1f} s)")
    ctx = bpx.Context(0)
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    work = db.num_residues * len(qs)

    # the cost model: trained on 100 seeded databases drawn like this one
    t = time.perf_counter()
    sample = costsched.sample_queries(qs, 16)
    model, counts, losses = costsched.fit(ctx, SKEWED_SPEC, sample, batches=100, subjects=2000, epochs=8)
    train_s = time.perf_counter() - t
    log(f"  cost model: trained in {train_s:.1f} s, epoch losses {[round(float(x), 4) for x in losses]}")

    gr, gq, launch, _ = bpx.autotune_batch(ctx, qs, ddb, 0, db.num_subjects // 10)
    b = ctx.prepare_batch(qs, group_residues=gr, group_queries=gq)
    groups = b.info()["groups"]
    log(f"  autotune: group_residues={gr} group_queries={gq} launch={launch}; {groups} groups")
    t = time.perf_counter()
    X = bpx.db_features(ctx, ddb, counts)
    cost_model = costsched.predicted_cost(model, X, db.lengths, len(qs), len(sample), groups)
    infer_s = time.perf_counter() - t

    def timed(cost, label):
        bpx.db_set_cost(ddb, cost)
        g = bpx.scan_batch(ctx, b, ddb, launch=launch, graph=True).wait()
        for _ in range(args.warmup):
            g.replay().wait()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.replay().wait()
        ms = (time.perf_counter() - t0) * 1000.0 / args.steps
        n = g.summary()[0]
        kern = g.timing()["kernel_ms"]
        g.close()
        log(f"  {label}: {ms:.1f} ms/step -> {work / ms / 1e6:.2f} G res/s")
        return ms, kern, n

    ms_len_default, _, n0 = timed(None, "static longest-first order")
    ms_model, kern, n2 = timed(cost_model, "learned-cost order")
    if n0 != n2:
        raise SystemExit(f"record counts differ between orders: {n0} {n2}")
    value = work / (ms_model / 1000.0)
    bytes_alg = groups * (db.num_residues + 12 * db.num_subjects) + 12 * db.num_subjects * len(qs) + 24 * n2
    achieved = bytes_alg / (kern / 1000.0) / 1e9
    out = {"metric": METRIC, "value": round(value, 1), "unit": "DB residues/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_model, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int8/int32",
           "data": "synthetic (seeded generator: skewed lengths, 20 % biased composition with tandem repeats)",
           "config": {"workload": "batched + learned-cost order: 1,024 queries vs skewed 500,000-subject DB",
                      "queries": len(qs), "subjects": db.num_subjects, "residues": db.num_residues,
                      "db_sha256": db.sha256, "records": int(n2), "groups": groups, "group_residues": gr,
                      "group_queries": gq, "launch": launch,
                      "ms_longest_first": round(ms_len_default, 3),
                      "gain_over_longest_first": round(ms_len_default / ms_model, 4),
                      "cost_model_train_s": round(train_s, 2), "cost_model_inference_s": round(infer_s, 3)},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                        "kernel": "k_scan (the graph of the batched scan)", "kernel_ms_mean": round(kern, 3),
                        "bytes_alg": int(bytes_alg), "passes": groups},
           "cpu_baseline": None}
    model.close()
    print(json.dumps(out), flush=True)


def serving(args):
    import threading
    import bpx
    log("generating the 2,000,000-subject database and the request stream ...")
    q0, db = plain_workload(args.threads)
    ctx = bpx.Context(0)
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    rates = [float(x) for x in args.rates.split(",")]
    n = args.requests
    lens = bpx.synth_query_lengths(5, n * len(rates), "lognormal", 20, 3000, 350, 0.6)
    queries = [bpx.synth_query(5, i, int(L)) for i, L in enumerate(lens)]
    # warm-up: one request through a server (code objects, pools)
    srv = bpx.Server(ctx, ddb)
    t = srv.submit(queries[0])
    info = srv.wait(t)
    srv.records(t, info["num_records"])
    lines = []
    all_results = {}
    for ri, rate in enumerate(rates):
        qs = queries[ri * n:(ri + 1) * n]
        rng = np.random.default_rng(100 + ri)
        gaps = rng.exponential(1.0 / rate, n)
        tickets = [None] * n
        t_start = time.perf_counter()

        def producer():
            t_next = time.perf_counter()
            for i, q in enumerate(qs):
                t_next += gaps[i]
                d = t_next - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets[i] = srv.submit(q)

        th = threading.Thread(target=producer)
        th.start()
        lat, sizes = [], []
        for i in range(n):
            while tickets[i] is None:
                time.sleep(0.0005)
            info = srv.wait(tickets[i])
            recs, subs = srv.records(tickets[i], info["num_records"])
            all_results[ri * n + i] = (recs, subs, info["stats"])
            lat.append(info["latency_ms"])
            sizes.append(info["batch_size"])
        th.join()
        wall = time.perf_counter() - t_start
        thr = db.num_residues * n / wall
        p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
        log(f"  rate {rate:g}/s: {n} requests in {wall:.2f} s -> {thr / 1e9:.2f} G res/s; latency p50 {p50:.1f} ms, "
            f"p99 {p99:.1f} ms; mean batch {np.mean(sizes):.1f}")
        lines.append({"rate_qps": rate, "requests": n, "seconds": round(wall, 3), "residues_per_s": round(thr, 1),
                      "p50_ms": round(p50, 2), "p99_ms": round(p99, 2), "mean_batch": round(float(np.mean(sizes)), 2)})
    c = srv.counters()
    srv.close()
    # every served result against the plain single-query scan
    mismatches = 0
    for k, (recs, subs, st) in all_results.items():
        dq = ctx.prepare_query(queries[k])
        r = bpx.scan(ctx, dq, ddb, subject_stats=False).wait()
        pr, ps = r.records()
        pst = r.summary()[1]
        r.close()
        dq.close()
        if not (np.array_equal(pr.view(np.int32), recs.view(np.int32)) and np.array_equal(ps, subs) and pst == st):
            mismatches += 1
    log(f"  {len(all_results)} served results checked against the plain scan: {mismatches} mismatches; {c}")
    if mismatches:
        raise SystemExit("served results differ from the plain scan")
    head = lines[min(1, len(lines) - 1)]
    out = {"metric": METRIC, "value": head["residues_per_s"], "unit": "DB residues/s", "n_gpus": 1,
           "steps": n, "warmup": 1, "ms_per_step": round(head["seconds"] * 1000.0 / n, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int8/int32",
           "data": "synthetic (seeded generator); Poisson arrivals",
           "config": {"workload": f"serving: Poisson single queries at {head['rate_qps']:g}/s vs the resident "
                                  "2,000,000-subject DB",
                      "p50_ms": head["p50_ms"], "p99_ms": head["p99_ms"], "rates": lines,
                      "query_lengths": "lognormal median 350 sigma 0.6 clipped 20-3000",
                      "checked_against_plain": len(all_results), "mismatches": mismatches,
                      "batches": c["batches"]},
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark of the gfx950 blastp word finder (two-hit word finding +
ungapped X-drop extension) on one MI355X.  Prints ONE JSON line; detail goes
to stderr.  Inputs are resident in HBM before any timed region (database
arena and query tables uploaded once).

--workload headline (default): the batched, cost-ordered scan of 1,024
queries (log-uniform 50-2,000 aa) against the 500,000-subject database
(bpx.workloads 3).  One step = one replay of the HIP graph that holds the
whole batched scan (counter resets, every launch class, compaction, copies)
plus the copy of its records to the host.  `value` is DB residues/s summed
over the 1,024 queries.  Beside it, in `config`:
  * the plain form -- each query scanned on its own (bpx_scan), the way the
    reference launches one query at a time -- and batched_over_plain;
  * the same batch under the static longest-first order, and the learned
    cost order's gain over it;
  * split_by_length: batched vs each-query-alone kernel time for the
    queries up to 1,000 aa (packed into groups) and the longer ones (groups
    of one), which shows where batched_over_plain is set;
  * the plain 500-aa query against the 2,000,000-subject database
    (bpx.workloads 2).
`roofline` is that of the batched step's k_scan launches (every launch
class), timed by HIP events around each graph replay on its stream (the
replay is k_scan but for a zero-fill, the compaction and small copies,
under 0.1 % of it by the rocprofv3 trace).

--workload plain: the 500-aa query against the 2,000,000-subject database
alone, one step = bpx_scan + bpx_result_wait.
--workload skewed: the batch against the skewed 500,000-subject database
(bpx.workloads 4), learned cost order vs longest-first.
--workload serving: Poisson single queries against the resident
2,000,000-subject database through the continuous-batching server; p50/p99
latency per arrival rate, every result checked against the plain scan.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "blastp-cuda_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "DB residues/s on 1 GPU, batched+binned vs plain; HBM GB/s; p50/p99 latency"
DATA = "synthetic (seeded generator, BLOSUM62 background residues, planted homologs)"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """Threads the CPU baseline may use: every CPU this process may run on."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def kernel_source_sha():
    """SHA-256 of the scan kernel's source: a measured traffic figure is used
    only while the kernel it was measured on is unchanged."""
    with open(os.path.join(ROOT, "blastp-cuda_amd", "csrc", "kernels.hip"), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def same_policy(a, b):
    """Packing policies compared field by field (a missing field is 0)."""
    keys = ("group_residues", "group_queries", "group_diag_slots")
    return a is not None and all(int(a.get(k, 0)) == int(b.get(k, 0)) for k in keys)


def parse_policy(text):
    """--policy GR,GQ,QCAP,SBUF,CLAIMS[,DIAG] -> (policy, launch)."""
    f = [int(x) for x in text.split(",")]
    policy = {"group_residues": f[0], "group_queries": f[1], "group_diag_slots": f[5] if len(f) > 5 else 0}
    return policy, {"queue_capacity": f[2], "subject_buffer": f[3], "claims_per_share": f[4]}


def measured_traffic(name, policy=None):
    """HBM bytes per step from a committed rocprofv3 --pmc summary
    (profiles/<name>.json), or (None, reason).  A batched figure holds for
    the packing policy (residues and queries per group) it was measured
    under only."""
    path = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(path):
        return None, "not measured"
    with open(path) as f:
        t = json.load(f)
    if t.get("kernels_hip_sha256") != kernel_source_sha():
        return None, f"profiles/{name}.json was measured on another kernel source"
    if policy is not None and not same_policy(t.get("policy"), policy):
        return None, f"profiles/{name}.json was measured under packing policy {t.get('policy')}, not {policy}"
    return t.get("bytes_per_launch"), f"profiles/{name}.json (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)"


# Vector-issue peak: 256 CUs x 4 SIMDs, one wave64 vector instruction per
# SIMD every 4 cycles (16 lanes wide), at the 2.4 GHz peak engine clock.
VALU_PEAK_G = 256 * 4 * 2.4 / 4


def issue_roofline(name, kernel_ms, policy=None):
    """The vector-issue roofline of the scan kernel: the vector instructions
    per launch from a committed rocprofv3 SQ_INSTS_VALU summary
    (profiles/<name>.json, used only while the kernel source is the one it
    was measured on) over the kernel time of this run, against the peak
    issue rate of the 1,024 SIMDs."""
    path = os.path.join(ROOT, "profiles", name + ".json")
    out = {"bound": "vector issue", "unit": "G wave64 VALU instructions/s", "peak": VALU_PEAK_G}
    if not os.path.exists(path):
        return dict(out, achieved=None, frac=None, source="not measured")
    with open(path) as f:
        t = json.load(f)
    if t.get("kernels_hip_sha256") != kernel_source_sha():
        return dict(out, achieved=None, frac=None, source=f"profiles/{name}.json was measured on another kernel source")
    if policy is not None and not same_policy(t.get("policy"), policy):
        return dict(out, achieved=None, frac=None,
                    source=f"profiles/{name}.json was measured under packing policy {t.get('policy')}, not {policy}")
    achieved = t["valu_per_launch"] / (kernel_ms / 1e3) / 1e9
    return dict(out, achieved=round(achieved, 1), frac=round(achieved / VALU_PEAK_G, 4),
                valu_per_launch=t["valu_per_launch"], source=f"profiles/{name}.json (rocprofv3 --pmc SQ_INSTS_VALU)")


def phase_summary(ph):
    """Per-wavefront phase times of one scan (bpx_result_phases, measured
    outside the timed steps).  The subject-streaming phase -- descriptor
    reads and subject residues copied from HBM into LDS -- takes
    `stage_share` of the wavefronts' time; applied to the launch span that is
    `stream_phase_ms`, over which it moves `bytes_staged`: its rate against
    the HBM peak.  `flush_share` is the share of the wavefronts' time in
    extension flushes; 1 - `busy_fraction` the share of the launches'
    wavefront slots left idle by wavefronts that finished before the
    slowest (what any reordering of the work could at most recover)."""
    share = ph["stage_ms_mean"] / ph["wave_ms_mean"] if ph["wave_ms_mean"] else 0.0
    ms = share * ph["span_ms"]
    gbs = ph["bytes_staged"] / (ms / 1e3) / 1e9 if ms else 0.0
    return {"stage_share": round(share, 4), "stream_phase_ms": round(ms, 3), "bytes_staged": int(ph["bytes_staged"]),
            "subjects_staged": int(ph["subjects_staged"]), "stream_phase_gbs": round(gbs, 1),
            "stream_phase_frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
            "flush_share": round(ph["flush_ms_mean"] / ph["wave_ms_mean"], 4) if ph["wave_ms_mean"] else 0.0,
            "busy_fraction": round(ph["busy_fraction"], 5), "span_ms": round(ph["span_ms"], 3),
            "launch_classes": int(ph["classes"]), "waves": int(ph["waves"])}


def median_rate(fn, work, runs):
    """work / median wall time of `runs` calls of fn (after one warm-up)."""
    fn()
    times = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t)
    return work / float(np.median(times)), times


# ---- CPU baselines (the C oracle; test infrastructure used as the checker's timing) ----

def cpu_baseline_plain(q, db, threads, budget_s=20.0):
    """The C oracle (same semantics, same inputs) on bounded prefixes of the
    database: on `threads` threads the median of 5 runs after a warm-up, on
    one thread the median of 3."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    lk = O.OracleLookup(q)

    def timed(nthreads, share, runs):
        probe = 2000
        t = time.perf_counter()
        O.run_db(q, db.residues, db.offsets[: probe + 1], db.lengths[:probe], nthreads=nthreads, lookup=lk)
        dt = max(time.perf_counter() - t, 1e-4)
        target = int(db.offsets[probe] / dt * budget_s * share / (runs + 1))
        n = int(np.searchsorted(db.offsets, min(target, int(db.offsets[-1]))))
        n = max(probe, min(n, db.num_subjects))
        res = int(db.offsets[n])
        rate, _ = median_rate(lambda: O.run_db(q, db.residues, db.offsets[: n + 1], db.lengths[:n],
                                               nthreads=nthreads, lookup=lk), res, runs)
        return n, res, rate

    n, res, value = timed(threads, 0.6, 5)
    n1, res1, value1 = timed(1, 0.4, 3)
    lk.close()
    return {"value": value, "unit": "DB residues/s", "cores": threads, "kind": "port",
            "host_cpus": host_cores(), "cpu_model": host_cpu_model(),
            "sample": f"first {n} subjects ({res} residues) of the same database and query; "
                      f"median of 5 runs after a warm-up, {threads} threads",
            "value_1thread": value1,
            "sample_1thread": f"first {n1} subjects ({res1} residues); median of 3 runs after a warm-up, 1 thread"}


def cpu_baseline_batched(qs, db, threads, budget_s=24.0):
    """The C oracle over every 16th query of the batch against a bounded
    prefix of the database (DB residues summed over the sampled queries per
    second): on `threads` threads the median of 5 runs after a warm-up, on
    one thread the median of 3 on a smaller prefix."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    sample_q = qs[::16]
    lks = [O.OracleLookup(q) for q in sample_q]

    def run(n, nthreads):
        for q, lk in zip(sample_q, lks):
            O.run_db(q, db.residues, db.offsets[: n + 1], db.lengths[:n], nthreads=nthreads, lookup=lk)

    def timed(nthreads, share, runs):
        # the oracle's threads claim 4,096-subject chunks (NCBI's OID chunk
        # iterator): a sample of fewer chunks than threads would leave
        # threads idle, so it holds at least one chunk per thread
        probe = 300
        t = time.perf_counter()
        run(probe, nthreads)
        dt = max(time.perf_counter() - t, 1e-4)
        n = int(max(4096 * nthreads, probe * budget_s * share / (runs + 1) / dt))
        n = min(db.num_subjects, n)
        work = int(db.offsets[n]) * len(sample_q)
        rate, _ = median_rate(lambda: run(n, nthreads), work, runs)
        return n, rate

    n, value = timed(threads, 0.65, 5)
    n1, value1 = timed(1, 0.35, 3)
    for lk in lks:
        lk.close()
    return {"value": value, "unit": "DB residues/s", "cores": threads, "kind": "port",
            "host_cpus": host_cores(), "cpu_model": host_cpu_model(),
            "sample": f"{len(sample_q)} of the 1,024 queries (every 16th) against the first {n} subjects; "
                      f"median of 5 runs after a warm-up, {threads} threads",
            "value_1thread": value1,
            "sample_1thread": f"the same queries against the first {n1} subjects; median of 3 runs after a "
                              "warm-up, 1 thread"}


# ---- the plain 500-aa x 2,000,000-subject scan ----------------------------------------

def plain_leg(ctx, threads, steps, warmup):
    """Generates and loads workload 2, then times `steps` plain scans (each
    bpx_scan + bpx_result_wait with per-subject counters)."""
    import bpx
    from bpx import workloads as W
    t0 = time.perf_counter()
    q, db = W.config2(threads)
    if db.sha256 != W.SHA256[2]:
        raise SystemExit(f"workload 2 database differs from the pinned one: {db.sha256}")
    log(f"  workload 2: {db.num_subjects} subjects, {db.num_residues} residues ({time.perf_counter() - t0:.1f} s)")
    t_load = time.perf_counter()
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    load_s = time.perf_counter() - t_load
    dq = ctx.prepare_query(q)

    def step():
        r = bpx.scan(ctx, dq, ddb, subject_stats=True).wait()
        t = r.timing()
        n, st = r.summary()
        r.close()
        return t, n, st

    for _ in range(warmup):
        step()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t, nrec, ust = step()
        kernel_ms.append(t["kernel_ms"])
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    kern = float(np.mean(kernel_ms))
    # per-wavefront phase times of one more scan, outside the timed steps
    r = bpx.scan(ctx, dq, ddb, subject_stats=True, phases=True).wait()
    phases = phase_summary(r.phases())
    r.close()
    # algorithmic HBM bytes of the launch: residues (1 B) + descriptors
    # (12 B) read, records (24 B) written; the per-subject counters (12 B)
    # go to pinned host memory
    bytes_alg = db.num_residues + 12 * db.num_subjects + 24 * nrec
    out = dict(q=q, db=db, ddb=ddb, dq=dq, ms=ms, kern=kern, kernel_ms=kernel_ms, nrec=nrec, ust=ust,
               load_s=load_s, bytes_alg=bytes_alg, value=db.num_residues / (ms / 1000.0), phases=phases)
    return out


def plain(args):
    import bpx
    log("the 500-aa x 2,000,000-subject workload ...")
    ctx = bpx.Context(0)
    P = plain_leg(ctx, args.threads, args.steps, args.warmup)
    db = P["db"]
    traffic, tsrc = measured_traffic("traffic_k_scan_plain_500aa_2M_subjects")
    achieved = P["bytes_alg"] / (P["kern"] / 1000.0) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": tsrc,
                "kernel": "k_scan<3,true,false,true>", "kernel_ms_mean": round(P["kern"], 4),
                "kernel_ms_p50": round(float(np.percentile(P["kernel_ms"], 50)), 4),
                "kernel_ms_p99": round(float(np.percentile(P["kernel_ms"], 99)), 4),
                "bytes_alg": int(P["bytes_alg"])}
    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU oracle on a bounded sample ...")
        cpu = cpu_baseline_plain(P["q"], db, args.threads)
    out = {"metric": METRIC, "value": round(P["value"], 1), "unit": "DB residues/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(P["ms"], 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int8/int32", "data": DATA,
           "config": {"workload": "plain: 1 query x 500 aa vs 2,000,000-subject DB",
                      "queries": 1, "query_len": 500, "subjects": db.num_subjects, "residues": db.num_residues,
                      "db_sha256": db.sha256, "records": int(P["nrec"]), "lookup_hits": P["ust"]["lookup_hits"],
                      "init_extends": P["ust"]["init_extends"], "db_load_s": round(P["load_s"], 4),
                      "cold_search_residues_per_s": round(db.num_residues / (P["load_s"] + P["ms"] / 1000.0), 1),
                      "phases": P["phases"]},
           "roofline": roofline,
           "roofline_issue": issue_roofline("sq_counters_k_scan_plain_500aa_2M_subjects", P["kern"]),
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


# ---- the headline: batched + cost-ordered vs plain on the 1,024-query workload ----------

def headline(args):
    import bpx
    from bpx import costsched
    from bpx import workloads as W
    log("workload 3: 1,024 queries (log-uniform 50-2,000 aa) x 500,000 subjects ...")
    t0 = time.perf_counter()
    qs, db = W.config3(args.threads)
    if db.sha256 != W.SHA256[3]:
        raise SystemExit(f"workload 3 database differs from the pinned one: {db.sha256}")
    nq = len(qs)
    qres = int(sum(len(q) for q in qs))
    log(f"  {nq} queries ({qres} residues), {db.num_subjects} subjects, {db.num_residues} residues "
        f"({time.perf_counter() - t0:.1f} s)")
    ctx = bpx.Context(0)
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    work = db.num_residues * nq  # DB residues summed over queries

    # the learned cost model: trained on 100 seeded databases drawn like this one
    t = time.perf_counter()
    sample = costsched.sample_queries(qs, 16)
    model, counts, losses = costsched.fit(ctx, W.CONFIG3_SPEC, sample, batches=100, subjects=2000, epochs=8)
    train_s = time.perf_counter() - t
    log(f"  cost model trained in {train_s:.1f} s; epoch losses {[round(float(x), 4) for x in losses]}")

    # packing policy and launch shape: autotuned on the whole database (a
    # group pass's fixed costs weigh more on a sample and skew the choice),
    # or given with --policy
    t = time.perf_counter()
    if args.policy:
        policy, launch = parse_policy(args.policy)
        log(f"  policy given: {policy} launch={launch}")
    else:
        policy, launch, trials = bpx.autotune_batch(ctx, qs, ddb, 0, db.num_subjects)
        log(f"  autotune: {policy} launch={launch} ({time.perf_counter() - t:.1f} s, "
            f"{len(trials)} trials, each printed above as it finished)")
    b = ctx.prepare_batch(qs, **policy)
    groups = b.info()["groups"]
    t = time.perf_counter()
    X = bpx.db_features(ctx, ddb, counts)
    cost = costsched.predicted_cost(model, X, db.lengths, nq, len(sample), groups)
    infer_s = time.perf_counter() - t
    model.close()

    def graph_steps(cost_order, steps, warmup):
        bpx.db_set_cost(ddb, cost_order)
        if args.eager:
            kms, n = [], 0
            t0 = time.perf_counter()
            for i in range(warmup + steps):
                if i == warmup:
                    kms, t0 = [], time.perf_counter()
                r = bpx.scan_batch(ctx, b, ddb, launch=launch).wait()
                kms.append(r.timing()["kernel_ms"])
                n = r.summary()[0]
                r.close()
            return (time.perf_counter() - t0) * 1000.0 / steps, kms, n
        g = bpx.scan_batch(ctx, b, ddb, launch=launch, graph=True).wait()
        for _ in range(warmup):
            g.replay().wait()
        kms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            g.replay().wait()
            kms.append(g.timing()["kernel_ms"])
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        n = g.summary()[0]
        g.close()
        return ms, kms, n

    # the static longest-first order (the reference's length ordering), then
    # the learned order: the headline, timed over exactly `steps` replays
    lf_steps = min(args.steps, 5)
    ms_lf = None
    if not args.only_batched:
        ms_lf, _, n_lf = graph_steps(None, lf_steps, 1)
        log(f"  longest-first order: {ms_lf:.1f} ms/step -> {work / ms_lf / 1e6:.2f} G res/s")
    ms, kernel_ms, nrec = graph_steps(cost, args.steps, args.warmup)
    value = work / (ms / 1000.0)
    log(f"  learned cost order: {ms:.1f} ms/step -> {value / 1e9:.2f} G res/s (graph device time "
        f"{np.mean(kernel_ms):.1f} ms)")
    if ms_lf is not None and n_lf != nrec:
        raise SystemExit(f"record counts differ between orders: {n_lf} vs {nrec}")
    # per-wavefront phase times of one more (eager) scan in the learned order
    r = bpx.scan_batch(ctx, b, ddb, launch=launch, phases=True).wait()
    batched_phases = phase_summary(r.phases())
    r.close()
    log(f"  batched phases: {batched_phases}")

    if args.only_batched:
        print(json.dumps({"workload": "batched only (profiling run)", "ms_per_step": round(ms, 3),
                          "value": round(value, 1), "kernel_ms_mean": round(float(np.mean(kernel_ms)), 3),
                          "records": int(nrec), "groups": groups}), flush=True)
        return
    # plain: every query on its own, as the reference launches one query at a
    # time (one warm-up pass over 32 queries, then the median of 3 passes)
    bpx.db_set_cost(ddb, None)
    dqs = [ctx.prepare_query(q) for q in qs]
    plain_records = [0]

    def plain_pass(sub=None):
        n = 0
        for dq in (dqs if sub is None else dqs[:sub]):
            r = bpx.scan(ctx, dq, ddb, subject_stats=False).wait()
            n += r.summary()[0]
            r.close()
        plain_records[0] = n

    plain_pass(32)
    plain_value, plain_times = median_rate(plain_pass, work, 3)
    if plain_records[0] != nrec:
        raise SystemExit(f"batched and plain record counts differ: {nrec} vs {plain_records[0]}")
    # where the ratio is set, outside the timed passes: each query's kernel
    # time alone, and one batched scan (longest-first order, the tuned policy
    # and launch) of the queries up to 1,000 aa and one of the longer ones
    alone_ms = np.zeros(len(dqs))
    for i, dq in enumerate(dqs):
        r = bpx.scan(ctx, dq, ddb, subject_stats=False).wait()
        alone_ms[i] = r.timing()["kernel_ms"]
        r.close()
    for dq in dqs:
        dq.close()
    split = {}
    for label, pick in (("queries_up_to_1000_aa", lambda L: L <= 1000), ("queries_over_1000_aa", lambda L: L > 1000)):
        idx = [i for i, q in enumerate(qs) if pick(len(q))]
        if not idx:
            continue
        bs = ctx.prepare_batch([qs[i] for i in idx], **policy)
        kb = []
        for _ in range(2):
            r = bpx.scan_batch(ctx, bs, ddb, launch=launch).wait()
            kb.append(r.timing()["kernel_ms"])
            r.close()
        ng = bs.info()["groups"]
        bs.close()
        alone = float(alone_ms[idx].sum())
        split[label] = {"queries": len(idx), "groups": ng, "batched_kernel_ms": round(min(kb), 1),
                        "alone_kernel_ms": round(alone, 1), "ratio": round(alone / min(kb), 3)}
    log(f"  plain (each query alone): {np.median(plain_times):.2f} s -> {plain_value / 1e9:.2f} G res/s; "
        f"batched over plain {value / plain_value:.3f}; by query length (kernel ms): {split}")
    b.close()
    ddb.close()

    # the plain 500-aa query against the 2,000,000-subject database
    P = plain_leg(ctx, args.threads, 20, 3)
    log(f"  workload 2 plain: {P['ms']:.2f} ms/step -> {P['value'] / 1e9:.2f} G res/s (k_scan {P['kern']:.3f} ms)")
    plain2 = {"workload": "plain: 1 query x 500 aa vs 2,000,000-subject DB", "value": round(P["value"], 1),
              "ms_per_step": round(P["ms"], 4), "steps": 20, "records": int(P["nrec"]),
              "db_sha256": P["db"].sha256, "kernel_ms_mean": round(P["kern"], 4),
              "roofline_frac": round(P["bytes_alg"] / (P["kern"] / 1000.0) / 1e9 / HBM_PEAK_GBS, 5),
              "db_load_s": round(P["load_s"], 4),
              "cold_search_residues_per_s": round(P["db"].num_residues / (P["load_s"] + P["ms"] / 1000.0), 1),
              "phases": P["phases"],
              "roofline_issue": issue_roofline("sq_counters_k_scan_plain_500aa_2M_subjects", P["kern"])}
    P["dq"].close()
    P["ddb"].close()

    # roofline of the batched step's k_scan launches (every launch class):
    # each group pass reads the residues and the 12-byte descriptors, and the
    # records (24 B) are written
    kern = float(np.mean(kernel_ms))
    bytes_alg = groups * (db.num_residues + 12 * db.num_subjects) + 24 * nrec
    achieved = bytes_alg / (kern / 1000.0) / 1e9
    traffic, tsrc = measured_traffic("traffic_k_scan_batched_1024_queries_500k_subjects",
                                   policy)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic, "traffic_source": tsrc,
                "kernel": "k_scan (all launch classes of one batched step)",
                "kernel_ms_mean": round(kern, 3),
                "kernel_ms_p50": round(float(np.percentile(kernel_ms, 50)), 3),
                "kernel_ms_p99": round(float(np.percentile(kernel_ms, 99)), 3),
                "bytes_alg": int(bytes_alg), "passes": groups}
    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU oracle on a bounded sample ...")
        cpu = cpu_baseline_batched(qs, db, args.threads)
    out = {"metric": METRIC, "value": round(value, 1), "unit": "DB residues/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int8/int32", "data": DATA,
           "config": {"workload": "batched + learned cost order: 1,024 queries (log-uniform 50-2,000 aa) vs "
                                  "500,000-subject DB",
                      "queries": nq, "query_residues": qres, "subjects": db.num_subjects,
                      "residues": db.num_residues, "db_sha256": db.sha256, "records": int(nrec),
                      "groups": groups, "policy": policy, "launch": launch,
                      "plain_value": round(plain_value, 1),
                      "plain_seconds": round(float(np.median(plain_times)), 3),
                      "batched_over_plain": round(value / plain_value, 3),
                      "ms_longest_first": round(ms_lf, 3), "steps_longest_first": lf_steps,
                      "gain_over_longest_first": round(ms_lf / ms, 4),
                      "cost_model_train_s": round(train_s, 2), "cost_model_inference_s": round(infer_s, 3),
                      "split_by_length": split,
                      "phases": batched_phases, "plain_2M": plain2},
           "roofline": roofline,
           "roofline_issue": issue_roofline("sq_counters_k_scan_batched_1024_queries_500k_subjects", kern,
                                            policy),
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: headline 10, plain 20, skewed 2)")
    ap.add_argument("--warmup", type=int, default=None, help="warm-up steps (default: headline 1, plain 3)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--threads", type=int, default=host_cores(),
                    help="host threads of the CPU baseline and the generator (default: every CPU available)")
    ap.add_argument("--workload", choices=("headline", "batched", "plain", "skewed", "serving"), default="headline")
    ap.add_argument("--rates", default="25,50,200,2000", help="serving: arrival rates (queries/s), comma-separated")
    ap.add_argument("--requests", type=int, default=120, help="serving: requests per arrival rate")
    ap.add_argument("--policy", default="", metavar="GR,GQ,QCAP,SBUF,CLAIMS[,DIAG]",
                    help="headline and skewed: group residues, group queries, hit-ring capacity, subject buffer, "
                         "claims per share and (optional) group diagonal slots to use instead of autotuning")
    ap.add_argument("--only-batched", action="store_true",
                    help="headline: time the batched graph only (no longest-first, per-query or 2,000,000-subject "
                         "legs), for profiling runs; prints a reduced line")
    ap.add_argument("--eager", action="store_true",
                    help="with --only-batched: launch each step's kernels directly instead of replaying a graph "
                         "(the same kernels, for counter-collection runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="first run a plain elementwise copy (torch.neg over int32) reading and writing BYTES, "
                         "3 times: the known-size kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.workload == "batched":
        args.workload = "headline"
    if args.steps is None:
        args.steps = {"headline": 10, "skewed": 2}.get(args.workload, 20)
    if args.warmup is None:
        args.warmup = 1 if args.workload in ("headline", "skewed") else 3
    os.environ.setdefault("BPX_TUNE_LOG", "1")
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the word finder would shard "
              "by subject across GPUs, but the library and this benchmark drive a single GPU.", file=sys.stderr)
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

    return {"headline": headline, "plain": plain, "skewed": skewed, "serving": serving}[args.workload](args)


def skewed(args):
    import bpx
    from bpx import costsched
    from bpx import workloads as W
    log("generating the skewed-database workload ...")
    t0 = time.perf_counter()
    qs, db = W.config4(args.threads)
    log(f"  {len(qs)} queries, {db.num_subjects} subjects, {db.num_residues} residues, sha256 {db.sha256} "
        f"({time.perf_counter() - t0:.1f} s)")
    ctx = bpx.Context(0)
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    work = db.num_residues * len(qs)

    # the cost model: trained on 100 seeded databases drawn like this one
    t = time.perf_counter()
    sample = costsched.sample_queries(qs, 16)
    model, counts, losses = costsched.fit(ctx, W.CONFIG4_SPEC, sample, batches=100, subjects=2000, epochs=8)
    train_s = time.perf_counter() - t
    log(f"  cost model: trained in {train_s:.1f} s, epoch losses {[round(float(x), 4) for x in losses]}")

    if args.policy:
        policy, launch = parse_policy(args.policy)
    else:
        policy, launch, _ = bpx.autotune_batch(ctx, qs, ddb, 0, db.num_subjects)
    b = ctx.prepare_batch(qs, **policy)
    groups = b.info()["groups"]
    log(f"  autotune: {policy} launch={launch}; {groups} groups")
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
        r = bpx.scan_batch(ctx, b, ddb, launch=launch, phases=True).wait()
        ph = phase_summary(r.phases())
        r.close()
        log(f"  {label}: {ms:.1f} ms/step -> {work / ms / 1e6:.2f} G res/s; phases {ph}")
        return ms, kern, n, ph

    ms_len_default, _, n0, ph_len = timed(None, "static longest-first order")
    ms_model, kern, n2, ph_model = timed(cost_model, "learned-cost order")
    if n0 != n2:
        raise SystemExit(f"record counts differ between orders: {n0} {n2}")
    value = work / (ms_model / 1000.0)
    bytes_alg = groups * (db.num_residues + 12 * db.num_subjects) + 24 * n2
    achieved = bytes_alg / (kern / 1000.0) / 1e9
    out = {"metric": METRIC, "value": round(value, 1), "unit": "DB residues/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_model, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int8/int32",
           "data": "synthetic (seeded generator: skewed lengths, 20 % biased composition with tandem repeats)",
           "config": {"workload": "batched + learned-cost order: 1,024 queries vs skewed 500,000-subject DB",
                      "queries": len(qs), "subjects": db.num_subjects, "residues": db.num_residues,
                      "db_sha256": db.sha256, "records": int(n2), "groups": groups, "policy": policy,
                      "launch": launch,
                      "ms_longest_first": round(ms_len_default, 3),
                      "gain_over_longest_first": round(ms_len_default / ms_model, 4),
                      "phases_longest_first": ph_len, "phases_learned": ph_model,
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
    from bpx import workloads as W
    log("generating the 2,000,000-subject database and the request stream ...")
    q0, db = W.config2(args.threads)
    ctx = bpx.Context(0)
    ddb = ctx.load_db(db.residues, db.offsets[:-1], db.lengths)
    rates = [float(x) for x in args.rates.split(",")]
    n = args.requests
    queries = W.serving_queries(n * len(rates))
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

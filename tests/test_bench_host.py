"""bench.py's host logic on the CPU: the --policy parser, the policy match
and kernel-source gate that decide whether a committed rocprofv3 summary
may be reported with a run, the vector-issue roofline and the phase-time
arithmetic.  No GPU is touched: bench.py imports the library lazily."""
import importlib.util
import json
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_parse_policy_with_and_without_diag_slots(bench):
    pol, launch = bench.parse_policy("2000,16,256,2048,4,2048")
    assert pol == {"group_residues": 2000, "group_queries": 16, "group_diag_slots": 2048}
    assert launch == {"queue_capacity": 256, "subject_buffer": 2048, "claims_per_share": 4}
    pol, launch = bench.parse_policy("1200,8,512,1024,0")
    assert pol["group_diag_slots"] == 0 and pol["group_queries"] == 8
    assert launch["claims_per_share"] == 0


def test_same_policy_treats_missing_fields_as_zero(bench):
    a = {"group_residues": 2000, "group_queries": 8}
    assert bench.same_policy(a, {"group_residues": 2000, "group_queries": 8, "group_diag_slots": 0})
    assert not bench.same_policy(a, {"group_residues": 2000, "group_queries": 8, "group_diag_slots": 2048})
    assert not bench.same_policy(None, a)


def test_traffic_is_reported_only_for_the_measured_kernel_and_policy(bench, monkeypatch):
    name = "traffic_k_scan_batched_1024_queries_500k_subjects"
    with open(os.path.join(ROOT, "profiles", name + ".json")) as f:
        t = json.load(f)
    policy = t["policy"]
    if t["kernels_hip_sha256"] == bench.kernel_source_sha():
        got, src = bench.measured_traffic(name, policy)
        assert got == t["bytes_per_launch"] and "rocprofv3" in src
        other = dict(policy, group_queries=policy["group_queries"] + 1)
        got, src = bench.measured_traffic(name, other)
        assert got is None and "packing policy" in src
    # another kernel source: no figure, whatever the policy
    monkeypatch.setattr(bench, "kernel_source_sha", lambda: "0" * 64)
    got, src = bench.measured_traffic(name, policy)
    assert got is None and "another kernel source" in src
    got, src = bench.measured_traffic("no_such_profile")
    assert got is None and src == "not measured"


def test_issue_roofline_arithmetic(bench, monkeypatch):
    # 1,024 SIMDs, one wave64 vector instruction each per 4 cycles at 2.4 GHz
    assert bench.VALU_PEAK_G == pytest.approx(614.4)
    name = "sq_counters_k_scan_plain_500aa_2M_subjects"
    with open(os.path.join(ROOT, "profiles", name + ".json")) as f:
        t = json.load(f)
    monkeypatch.setattr(bench, "kernel_source_sha", lambda: t["kernels_hip_sha256"])
    r = bench.issue_roofline(name, 10.0)  # 10 ms per launch
    assert r["achieved"] == pytest.approx(t["valu_per_launch"] / 1e-2 / 1e9, abs=0.05)
    assert r["frac"] == pytest.approx(r["achieved"] / 614.4, abs=1e-4)
    monkeypatch.setattr(bench, "kernel_source_sha", lambda: "0" * 64)
    assert bench.issue_roofline(name, 10.0)["achieved"] is None


def test_phase_summary_arithmetic(bench):
    # wavefronts spend 1 of their 10 ms staging over a 20 ms span: the phase
    # is 2 ms of the launch, over which 8 GB move -> 4,000 GB/s, half of peak
    ph = {"stage_ms_mean": 1.0, "wave_ms_mean": 10.0, "span_ms": 20.0, "bytes_staged": 8_000_000_000,
          "subjects_staged": 1000, "flush_ms_mean": 2.5, "busy_fraction": 0.97, "classes": 2, "waves": 4096}
    s = bench.phase_summary(ph)
    assert s["stage_share"] == pytest.approx(0.1)
    assert s["stream_phase_ms"] == pytest.approx(2.0)
    assert s["stream_phase_gbs"] == pytest.approx(4000.0)
    assert s["stream_phase_frac_of_peak"] == pytest.approx(0.5)
    assert s["flush_share"] == pytest.approx(0.25)
    assert s["busy_fraction"] == pytest.approx(0.97)

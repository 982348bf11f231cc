"""The host-only library sources (lookup-table builder, BLOSUM62 table,
synthetic generator, NCBI volume and .gpu readers) and the CPU oracle, built
on the CPU with AddressSanitizer + UndefinedBehaviorSanitizer and, for the
threaded parts (the generator and the oracle's whole-database run), with
ThreadSanitizer, and driven by tests/sanitize_harness.cpp: table builds
against the oracle's brute force, a threaded database and the oracle's run
over it on one and four threads, a volume round trip, hostile index headers
and malformed .gpuinfo text.  Any sanitizer report fails the test."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "blastp-cuda_amd", "csrc")
HOST_SOURCES = [os.path.join(CSRC, f) for f in ("table.cpp", "blosum62.cpp", "synth.cpp", "volume.cpp")]
HARNESS = os.path.join(ROOT, "tests", "sanitize_harness.cpp")
ORACLE = os.path.join(ROOT, "oracle", "oracle.c")
INCLUDES = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-I" + os.path.join(ROOT, "oracle")]


def build_and_run(tmp_path, flags, env_extra):
    if shutil.which("g++") is None or shutil.which("gcc") is None:
        pytest.skip("gcc / g++ not available")
    objs = []
    oracle_obj = str(tmp_path / "oracle.o")
    subprocess.run(["gcc", "-std=c11", "-D_GNU_SOURCE", "-O1", "-g", "-fno-omit-frame-pointer", *flags, *INCLUDES,
                    "-c", ORACLE, "-o", oracle_obj], check=True)
    objs.append(oracle_obj)
    for src in HOST_SOURCES + [HARNESS]:
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", *flags, *INCLUDES,
                        "-c", src, "-o", obj], check=True)
        objs.append(obj)
    exe = str(tmp_path / "harness")
    subprocess.run(["g++", *flags, *objs, "-o", exe, "-lpthread"], check=True)
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    env = dict(os.environ, **env_extra)
    r = subprocess.run([exe, str(scratch)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, f"harness failed ({r.returncode}):\n{r.stdout}\n{r.stderr[-6000:]}"
    assert "all checks passed" in r.stdout
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-6000:]


def test_host_sources_and_oracle_under_address_and_undefined_sanitizers(tmp_path):
    build_and_run(tmp_path, ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
                  {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1"})


def test_threaded_generator_and_oracle_under_thread_sanitizer(tmp_path):
    build_and_run(tmp_path, ["-fsanitize=thread"], {"TSAN_OPTIONS": "halt_on_error=1"})

"""Host side of the product, checked on the CPU: the built-in scoring table,
the lookup-table builder (against the oracle's brute-force table), the
synthetic generator, and the C ABI's exported symbols."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

import bpx
import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STD20 = np.array([1, 16, 13, 4, 3, 15, 5, 7, 8, 9, 11, 10, 12, 6, 14, 17, 18, 20, 22, 19], dtype=np.uint8)


def test_product_matrix_equals_reference_table():
    with open(os.path.join(ROOT, "tests", "golden", "blosum62_ncbistdaa.json")) as f:
        ref = np.array(json.load(f)["rows"], dtype=np.int32)
    np.testing.assert_array_equal(bpx.blosum62().astype(np.int32), ref)


def _tables_equal(q, T=11, word=3, M=None):
    cs, ent = bpx.table_build(q, T, word, M)
    ocs, oent = O.OracleLookup(q, T, word, M).csr()
    np.testing.assert_array_equal(cs, ocs)
    np.testing.assert_array_equal(ent, oent)
    return cs, ent


@pytest.mark.parametrize("seed", range(8))
def test_table_matches_oracle_random(seed):
    rng = np.random.default_rng(seed)
    q = rng.choice(STD20, int(rng.integers(3, 700)))
    _tables_equal(q, T=int(rng.integers(9, 14)))


def test_table_matches_oracle_special():
    rng = np.random.default_rng(9)
    # nonstandard residues: X (21), B (2), Z (23), U (24), * (25), O (26), J (27)
    q = rng.choice(np.arange(1, 28, dtype=np.uint8), 300)
    _tables_equal(q)
    # low complexity: one word repeated -> a very large cell
    cs, ent = _tables_equal(np.full(500, 11, dtype=np.uint8))
    assert np.diff(cs).max() == 498
    # word length 2 and an asymmetric matrix
    M = rng.integers(-4, 7, (28, 28)).astype(np.int8)
    M[0, :] = -128
    M[:, 0] = -128
    _tables_equal(rng.choice(STD20, 120), T=8, word=2, M=M)
    _tables_equal(rng.choice(STD20, 120), T=9, word=3, M=M)
    # degenerate lengths
    for n in (0, 1, 2, 3):
        _tables_equal(rng.choice(STD20, n))


def test_table_rejects_bad_input():
    with pytest.raises(bpx.BpxError):
        bpx.table_build(np.array([1, 2, 30], np.uint8))
    with pytest.raises(bpx.BpxError):
        bpx.table_build(np.array([1, 2, 3], np.uint8), word_length=4)


def test_sha256_known_answers():
    assert bpx.sha256_hex(np.frombuffer(b"abc", np.uint8)) == \
        "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"
    assert bpx.sha256_hex(np.zeros(0, np.uint8)) == \
        "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"
    msg = b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq"
    assert bpx.sha256_hex(np.frombuffer(msg, np.uint8)) == \
        "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1"


def test_synth_deterministic_and_shaped():
    q = bpx.synth_query(0, 0, 300)
    assert len(q) == 300 and set(np.unique(q)) <= set(STD20.tolist())
    assert (bpx.synth_query(0, 0, 300) == q).all() and not (bpx.synth_query(1, 0, 300) == q).all()
    spec = dict(seed=5, num_subjects=2000, len_median=300, len_sigma=0.6, len_min=30, len_max=5000,
                homolog_count=20, identity_min=0.5, identity_max=0.5, indel_min=0.02, indel_max=0.02,
                biased_fraction=0.2, dirichlet_alpha=0.3, motif_min=5, motif_max=30)
    a = bpx.synth_db(spec, [q], nthreads=1)
    b = bpx.synth_db(spec, [q], nthreads=7)
    assert a.sha256 == b.sha256
    assert (a.residues == b.residues).all() and (a.lengths == b.lengths).all()
    assert a.num_subjects == 2000 and (a.kinds == 1).sum() == 20
    assert 0.1 < (a.kinds == 2).mean() < 0.3
    assert a.lengths.min() >= 30 and a.lengths.max() <= 5000
    assert (np.diff(a.offsets.astype(np.int64)) == a.lengths).all()
    med = float(np.median(a.lengths))
    assert 250 < med < 360
    spec2 = dict(spec, seed=6)
    assert bpx.synth_db(spec2, [q]).sha256 != a.sha256


def test_synth_query_lengths():
    ln = bpx.synth_query_lengths(3, 1024, "loguniform", 50, 2000)
    assert ln.min() >= 50 and ln.max() <= 2000
    assert 400 < ln.mean() < 650  # log-uniform 50..2000 has mean ~529
    ln2 = bpx.synth_query_lengths(3, 4000, "lognormal", 20, 3000, 350, 0.6)
    assert 320 < np.median(ln2) < 380


def _declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(bpx_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(bpx.LIB_PATH)
    names = _declared_functions()
    assert len(names) > 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_package_fails_loudly_without_the_library(tmp_path):
    """No CPU fallback: the binding refuses to import when libbpx.so is absent."""
    import shutil
    import subprocess
    import sys as _sys
    pkg = tmp_path / "bpx"
    pkg.mkdir()
    shutil.copy(os.path.join(os.path.dirname(bpx.__file__), "__init__.py"), pkg / "__init__.py")
    r = subprocess.run([_sys.executable, "-c", "import bpx"], cwd=str(tmp_path), capture_output=True, text=True)
    assert r.returncode != 0 and "ImportError" in r.stderr and "not built" in r.stderr


def test_workload_databases_are_pinned():
    """Each configured database is the same bytes on every machine and with
    any thread count: its SHA-256 (lengths, then residues) is pinned."""
    from bpx import workloads as W
    q1, d1 = W.config1(3)
    assert len(q1) == 300 and d1.num_subjects == 10_000 and d1.sha256 == W.SHA256[1]
    _, d3 = W.config3(8)
    assert d3.num_subjects == 500_000 and d3.sha256 == W.SHA256[3]
    del d3
    qs, d4 = W.config4(8)
    assert len(qs) == 1024 and d4.num_subjects == 500_000 and d4.sha256 == W.SHA256[4]
    assert d4.lengths.max() <= 35_000 and 0.15 < (d4.kinds == 2).mean() < 0.25
    del d4
    q2, d2 = W.config2(8)
    assert len(q2) == 500 and d2.num_subjects == 2_000_000 and d2.sha256 == W.SHA256[2]

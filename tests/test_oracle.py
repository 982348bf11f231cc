"""Pins the CPU oracle before anything is compared against it.

* the scoring table against the reference's own table (golden fixture read
  from gpu_blastp.c:46-75) and the published BLOSUM62 properties;
* the lookup table against a brute-force Python enumeration;
* the extensions against closed forms (dropoff = infinity: maximum
  suffix / prefix sums written with numpy cumsum/argmax);
* the two-hit scan against hand-derived known answers (tests/golden/
  known_answers.json, derivations in that file) and against the naive
  restatement in ref_straight.py on random and adversarial inputs,
  including tiny diagonal masks where diagonals alias;
* the per-subject-fresh table against NCBI's never-reset running offset.
"""
import json
import os

import numpy as np
import pytest

import pyoracle as O
import ref_straight as R

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
STD20 = np.array([1, 16, 13, 4, 3, 15, 5, 7, 8, 9, 11, 10, 12, 6, 14, 17, 18, 20, 22, 19], dtype=np.uint8)


def ref_table():
    with open(os.path.join(GOLDEN, "blosum62_ncbistdaa.json")) as f:
        return np.array(json.load(f)["rows"], dtype=np.int32)


def test_matrix_equals_reference_table():
    np.testing.assert_array_equal(O.blosum62().astype(np.int32), ref_table())


def test_matrix_published_properties():
    M = O.blosum62().astype(np.int32)
    assert (M == M.T).all()
    # a few published BLOSUM62 entries (ncbistdaa: A=1 C=3 W=20 Y=22 F=6 D=4 E=5 I=9 V=19)
    assert M[20, 20] == 11 and M[3, 3] == 9 and M[1, 1] == 4
    assert M[22, 6] == 3 and M[4, 5] == 2 and M[9, 19] == 3 and M[20, 1] == -3
    # every standard self-score is >= 4, so every standard 3-word scores >= 12 > T=11
    assert min(M[c, c] for c in STD20) == 4
    assert (M[0, :] == -128).all() and (M[:, 0] == -128).all()


def brute_table(query, M, T=11, word=3):
    cells = {}
    for p in range(len(query) - word + 1):
        qw = [int(x) for x in query[p:p + word]]
        for a in range(28):
            for b in range(28):
                for c in (range(28) if word == 3 else [0]):
                    w = (a, b, c)[:word]
                    sc = sum(int(M[qw[i]][w[i]]) for i in range(word))
                    if sc >= T or list(w) == qw:
                        idx = 0
                        for x in w:
                            idx = (idx << 5) | x
                        cells.setdefault(idx, []).append(p)
    return cells


@pytest.mark.parametrize("seed,qlen,T", [(1, 7, 11), (2, 12, 11), (3, 9, 13), (4, 6, 8)])
def test_lookup_matches_bruteforce(seed, qlen, T):
    rng = np.random.default_rng(seed)
    q = rng.choice(STD20, qlen)
    if seed == 3:
        q[2] = 21  # an X: its self-score (-3) is below T, so the exact word must still be listed
    M = O.blosum62()
    cs, ent = O.OracleLookup(q, T).csr()
    cells = brute_table(q, M, T)
    for idx in range(len(cs) - 1):
        got = list(ent[cs[idx]:cs[idx + 1]])
        assert got == cells.get(idx, []), idx
    assert int(cs[-1]) == sum(len(v) for v in cells.values())


def test_lookup_matches_bruteforce_x():
    test_lookup_matches_bruteforce(3, 9, 13)


def test_lookup_word2_matches_bruteforce():
    q = np.array([1, 5, 9, 11, 20, 22, 3], dtype=np.uint8)
    cs, ent = O.OracleLookup(q, 8, 2).csr()
    cells = brute_table(q, O.blosum62(), 8, 2)
    for idx in range(len(cs) - 1):
        assert list(ent[cs[idx]:cs[idx + 1]]) == cells.get(idx, [])


# ---- extensions: closed forms at dropoff = infinity ----------------------------
BIG = 1 << 30


def left_closed(m):
    """max(0, max prefix sum), length = first argmax + 1 (strict improvement)."""
    c = np.cumsum(m)
    if len(c) == 0 or c.max() <= 0:
        return 0, 0
    return int(c.max()), int(np.argmax(c)) + 1


def right_closed(m, init):
    r = init + np.cumsum(m)
    stop = np.nonzero(r <= 0)[0]
    end = int(stop[0]) if len(stop) else len(m)
    seen = r[: end + 1] if len(stop) else r
    if len(seen) == 0 or seen.max() <= init:
        best, length = init, 0
    else:
        best, length = int(seen.max()), int(np.argmax(seen)) + 1
    return best, length, end


@pytest.mark.parametrize("seed", range(12))
def test_extend_left_closed_form(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 40))
    s = rng.choice(STD20, n)
    q = rng.choice(STD20, n) if seed % 3 else s.copy()
    M = O.blosum62().astype(np.int64)
    s_off = q_off = n - 1
    scores = [M[q[q_off - k], s[s_off - k]] for k in range(n)]
    assert O.extend_left(s, q, s_off, q_off, BIG) == left_closed(np.array(scores))


@pytest.mark.parametrize("seed", range(12))
def test_extend_right_closed_form(seed):
    rng = np.random.default_rng(200 + seed)
    n = int(rng.integers(1, 40))
    s = rng.choice(STD20, n)
    q = rng.choice(STD20, n) if seed % 3 else s.copy()
    init = int(rng.integers(0, 20))
    M = O.blosum62().astype(np.int64)
    scores = np.array([M[q[k], s[k]] for k in range(n)])
    best, length, end = right_closed(scores, init)
    got = O.extend_right(s, q, 0, 0, BIG, init)
    assert got == (best, length, end)


def test_extend_xdrop_stops_at_drop():
    # identical prefix (+), then a long mismatching run: X-drop must stop
    # exactly when max - score >= dropoff
    M = O.blosum62().astype(np.int64)
    s = np.array([20, 20, 20, 14, 14, 14, 14, 14, 14], dtype=np.uint8)   # W W W P...
    q = np.array([20, 20, 20, 20, 20, 20, 20, 20, 20], dtype=np.uint8)   # W ...
    # right from 0: 11, 22, 33, then P vs W = -4 each: 29, 25, 21, 17 -> 33-17 = 16 >= 16 stop at k=6
    best, length, s_last = O.extend_right(s, q, 0, 0, 16, 0)
    assert (best, length, s_last) == (33, 3, 6)
    assert M[20, 14] == -4


# ---- two-hit scan ---------------------------------------------------------------

def identity_matrix(match=5, mismatch=-4):
    M = np.full((28, 28), mismatch, dtype=np.int8)
    np.fill_diagonal(M, match)
    M[0, :] = -128
    M[:, 0] = -128
    return M


def run_one(query, subject, M=None, **kw):
    prm = O.params(len(query), **kw)
    counters, recoff, recs = O.run_db(query, subject, np.array([0, len(subject)], np.uint64),
                                      np.array([len(subject)], np.uint32), prm, 1, M)
    return counters[0], recs


def test_known_answers():
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        cases = json.load(f)["cases"]
    import bpx_letters
    M = identity_matrix()
    for case in cases:
        q = bpx_letters.encode(case["query"])
        s = bpx_letters.encode(case["subject"])
        cnt, recs = run_one(q, s, M, **case.get("params", {}))
        assert list(cnt) == case["counters"], case["name"]
        got = [list(r) for r in recs.tolist()]
        assert got == case["records"], case["name"]


def _compare_straight(q, s, M, T=11, window=40, xdrop=16, cutoff=41, diag_len=None):
    dl = diag_len or O.diag_length(len(q), window)
    cnt, recs = run_one(q, s, M, threshold=T, window=window, x_dropoff=xdrop, cutoff=cutoff,
                        diag_mask=dl - 1)
    h, sv, ex, rr = R.scan(q, s, M, 3, T, window, xdrop, cutoff, dl)
    assert list(cnt) == [h, sv, ex]
    assert sorted(map(tuple, recs.tolist())) == sorted(rr)


@pytest.mark.parametrize("seed", range(20))
def test_scan_matches_straightline_random(seed):
    rng = np.random.default_rng(1000 + seed)
    M = O.blosum62()
    q = rng.choice(STD20, int(rng.integers(10, 60)))
    s = rng.choice(STD20, int(rng.integers(3, 120)))
    if seed % 2:  # plant a mutated copy of part of the query so extensions fire
        a = int(rng.integers(0, len(q) - 5))
        seg = q[a:].copy()
        mut = rng.random(len(seg)) < 0.3
        seg[mut] = rng.choice(STD20, int(mut.sum()))
        at = int(rng.integers(0, len(s)))
        s = np.concatenate([s[:at], seg, s[at:]])
    _compare_straight(q, s, M, cutoff=int(rng.integers(10, 45)))


@pytest.mark.parametrize("seed", range(6))
def test_scan_matches_straightline_asymmetric(seed):
    """A non-symmetric score table pins the orientation M[query][subject]
    (BLOSUM62 is symmetric and cannot)."""
    rng = np.random.default_rng(4000 + seed)
    M = rng.integers(-4, 7, (28, 28)).astype(np.int8)
    M[0, :] = -128
    M[:, 0] = -128
    q = rng.choice(STD20, int(rng.integers(10, 50)))
    s = rng.choice(STD20, int(rng.integers(10, 150)))
    s = np.concatenate([s, q[::-1], q])
    _compare_straight(q, s, M, T=9, cutoff=20)


@pytest.mark.parametrize("seed", range(10))
def test_scan_matches_straightline_aliasing(seed):
    """Diagonal tables far smaller than qlen + window: distinct diagonals share
    a slot, and the flag of one must gate hits on the other."""
    rng = np.random.default_rng(2000 + seed)
    M = O.blosum62()
    q = rng.choice(STD20[:6], int(rng.integers(8, 40)))   # low complexity: many hits
    s = rng.choice(STD20[:6], int(rng.integers(8, 150)))
    _compare_straight(q, s, M, window=int(rng.integers(4, 40)), diag_len=int(2 ** rng.integers(2, 6)),
                      cutoff=20)


@pytest.mark.parametrize("seed", range(6))
def test_scan_matches_straightline_repeats(seed):
    """Tandem repeats: the same word on many diagonals, runs on one diagonal."""
    rng = np.random.default_rng(3000 + seed)
    motif = rng.choice(STD20, int(rng.integers(2, 7)))
    q = np.tile(motif, 12)[: int(rng.integers(10, 50))]
    s = np.tile(motif, 30)[: int(rng.integers(10, 150))]
    _compare_straight(q, s, O.blosum62(), cutoff=25)


def test_edge_lengths():
    M = O.blosum62()
    q = np.array([1, 3, 4, 5], dtype=np.uint8)
    for s in (np.zeros(0, np.uint8), np.array([1], np.uint8), np.array([1, 3], np.uint8),
              np.array([1, 3, 4], np.uint8)):
        cnt, recs = run_one(q, s, M)
        h, sv, ex, rr = R.scan(q, s, M)
        assert list(cnt) == [h, sv, ex] and len(recs) == len(rr) == 0
    # query shorter than a word: no hits at all
    cnt, recs = run_one(np.array([1, 3], np.uint8), np.array([1, 3, 4, 5], np.uint8), M)
    assert list(cnt) == [0, 0, 0]


def test_fresh_table_equals_running_offset():
    """NCBI never clears the diagonal table between subjects; it moves the
    running offset by len + window.  Each subject must see the same results
    as with a fresh table."""
    rng = np.random.default_rng(7)
    q = rng.choice(STD20, 80)
    subjects = []
    for i in range(40):
        s = rng.choice(STD20, int(rng.integers(1, 200)))
        if i % 3 == 0:
            s = np.concatenate([s, q[10:60]])
        subjects.append(s)
    persistent, final_off = O.scan_subject_persistent(q, subjects)
    assert final_off == 40 + sum(len(s) + 40 for s in subjects)
    for s, (cnt, recs) in zip(subjects, persistent):
        c2, r2 = run_one(q, s)
        assert list(cnt) == list(c2)
        assert recs.tolist() == r2.tolist()


def test_record_invariants():
    rng = np.random.default_rng(11)
    M = O.blosum62().astype(np.int64)
    q = rng.choice(STD20, 200)
    subs = []
    for i in range(30):
        s = rng.choice(STD20, int(rng.integers(50, 400)))
        seg = q.copy()
        mut = rng.random(len(seg)) < 0.4
        seg[mut] = rng.choice(STD20, int(mut.sum()))
        at = int(rng.integers(0, len(s)))
        subs.append(np.concatenate([s[:at], seg, s[at:]]))
    lens = np.array([len(s) for s in subs], np.uint32)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    counters, recoff, recs = O.run_db(q, np.concatenate(subs), offs, lens, nthreads=3)
    assert len(recs) > 0
    for i, s in enumerate(subs):
        for r in recs[recoff[i]:recoff[i + 1]]:
            hq, hs, qo, so, ln, sc = (int(x) for x in r)
            assert hq - hs == qo - so
            assert 0 <= hs and hs + ln <= len(s) and 0 <= hq and hq + ln <= len(q)
            assert sc >= 41
            assert sc == sum(M[q[hq + k], s[hs + k]] for k in range(ln))
        assert counters[i, 1] == recoff[i + 1] - recoff[i]
        assert counters[i, 2] >= counters[i, 1]


def test_threads_do_not_change_results():
    rng = np.random.default_rng(5)
    q = rng.choice(STD20, 150)
    lens = rng.integers(1, 500, 3000).astype(np.uint32)
    res = rng.choice(STD20, int(lens.sum()))
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    a = O.run_db(q, res, offs, lens, nthreads=1)
    b = O.run_db(q, res, offs, lens, nthreads=4)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    assert a[2].tolist() == b[2].tolist()

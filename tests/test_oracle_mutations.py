"""The oracle's pins must catch plausible mistakes.  Each case compiles a
copy of oracle/oracle.c with one planted mistake (on the CPU, into a temp
dir), runs the known-answer and straight-line checks against it in a child
process, and expects them to fail."""
import os
import shutil
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "oracle.c")

MUTATIONS = [
    ("window boundary", "if (diff >= window) {", "if (diff > window) {"),
    ("word boundary", "if (diff < word)\n                continue;", "if (diff <= word)\n                continue;"),
    ("left boundary", "last + word, s, q, p->x_dropoff, word);", "last + word - 1, s, q, p->x_dropoff, word);"),
    ("flag reset point", "st->last_hit = x.s_last_off - (word - 1) + diag_offset;",
     "st->last_hit = x.s_last_off - word + diag_offset;"),
    ("right stop on score", "if (score <= 0 || maxscore - score >= dropoff)", "if (maxscore - score >= dropoff)"),
    ("xdrop comparison", "if (maxscore - score >= dropoff)\n            break;\n    }\n    or_ext_left",
     "if (maxscore - score > dropoff)\n            break;\n    }\n    or_ext_left"),
    ("transposed matrix", "score += M[query[q_off - n + i] * OR_ALPHABET + subject[s_off - n + i]];",
     "score += M[query[q_off - n + i] + subject[s_off - n + i] * OR_ALPHABET];"),
    # (s - q) & mask would be an equivalent mutant: negation mod 2^k keeps
    # exactly the same hits together, so it is not listed.
    ("diagonal formula", "const int32_t d = (q - s) & mask;", "const int32_t d = (q + s) & mask;"),
    ("mask dropped", "const int32_t d = (q - s) & mask;", "const int32_t d = (q - s) & (mask >> 1);"),
    ("prefix anchor", "right_d = i + 1;", "right_d = i;"),
    ("exact word dropped", "if (score >= threshold || idx == self) {", "if (score >= threshold) {"),
]

CHECK = textwrap.dedent("""
    import sys
    sys.path[:0] = [{tests!r}, {oracle!r}]
    import test_oracle as T
    fails = 0
    for fn in [T.test_known_answers, T.test_lookup_matches_bruteforce_x]:
        try:
            fn()
        except AssertionError:
            fails += 1
    for seed in range(20):
        try:
            T.test_scan_matches_straightline_random(seed)
        except AssertionError:
            fails += 1
    for seed in range(6):
        try:
            T.test_scan_matches_straightline_asymmetric(seed)
        except AssertionError:
            fails += 1
    for seed in range(10):
        try:
            T.test_scan_matches_straightline_aliasing(seed)
        except AssertionError:
            fails += 1
    for seed in range(12):
        for fn in (T.test_extend_left_closed_form, T.test_extend_right_closed_form):
            try:
                fn(seed)
            except AssertionError:
                fails += 1
    try:
        T.test_extend_xdrop_stops_at_drop()
    except AssertionError:
        fails += 1
    print("FAILS", fails)
""")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("name,old,new", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_planted_mistake_is_caught(tmp_path, name, old, new):
    src = open(SRC).read()
    assert src.count(old) == 1, f"mutation anchor not unique/present: {name}"
    mutated = tmp_path / "oracle.c"
    mutated.write_text(src.replace(old, new))
    shutil.copy(os.path.join(ROOT, "oracle", "oracle.h"), tmp_path / "oracle.h")
    lib = tmp_path / "liboracle.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-std=c11", "-D_GNU_SOURCE", "-o", str(lib),
                    str(mutated), "-lpthread"], check=True)
    env = dict(os.environ, BPX_ORACLE_LIB=str(lib))
    code = CHECK.format(tests=os.path.join(ROOT, "tests"), oracle=os.path.join(ROOT, "oracle"))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert "FAILS" in out.stdout, out.stderr[-2000:]
    fails = int(out.stdout.split("FAILS")[1].split()[0])
    assert fails > 0, f"planted mistake '{name}' went unnoticed"

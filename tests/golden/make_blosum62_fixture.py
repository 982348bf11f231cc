"""Writes blosum62_ncbistdaa.json: the 28x28 substitution table that the
reference stages on its GPU (gpu_blastp.c:46-75), read from that file as
data -- 784 integers in ncbistdaa order, rows = query residue.

Run in a container that has the reference tree mounted; the JSON it writes
is the committed fixture (the reference is not needed at test time).
"""
import json
import os
import re
import sys

SRC = os.environ.get("BPX_REFERENCE", "/root/reference") + "/gpu_blastp.c"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with open(SRC) as f:
        lines = f.read().split("\n")
    rows = []
    # the table body is lines 47-74 of gpu_blastp.c (1-based), one row a line
    for line in lines[46:74]:
        m = re.search(r"\{([^}]*)\}", line)
        rows.append([int(x) for x in m.group(1).split(",")])
    assert len(rows) == 28 and all(len(r) == 28 for r in rows)
    out = {
        "source": "gpu_blastp.c:46-75",
        "alphabet": "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ",
        "rows": rows,
    }
    with open(os.path.join(HERE, "blosum62_ncbistdaa.json"), "w") as f:
        json.dump(out, f, indent=None, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    sys.exit(main())

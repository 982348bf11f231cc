"""NCBI protein database volumes: the writer/reader round trip, the index
header read by an independent restatement of the reference's own header walk
(functions.cpp:181-215), alias DBLIST resolution (functions.cpp:217-353) and
malformed files.  Parity with volumes written by NCBI's makeblastdb is
unpinned here: no such file ships with the reference."""
import os
import struct

import numpy as np
import pytest

import bpx


def random_db(rng, n):
    lens = rng.integers(0, 400, n).astype(np.uint32)
    lens[::7] = 0
    offs = np.concatenate([[0], np.cumsum(lens.astype(np.uint64))]).astype(np.uint64)
    res = rng.integers(1, 26, int(offs[-1])).astype(np.uint8)
    return res, offs, lens


def reference_header_walk(buf):
    """functions.cpp:181-215: big-endian version, type, title, date, sequence count."""
    version, typ, tlen = struct.unpack(">III", buf[:12])
    title = buf[12:12 + tlen]
    (dlen,) = struct.unpack(">I", buf[12 + tlen:16 + tlen])
    (nseq,) = struct.unpack(">I", buf[16 + tlen + dlen:20 + tlen + dlen])
    total, = struct.unpack("<Q", buf[20 + tlen + dlen:28 + tlen + dlen])
    (mx,) = struct.unpack(">I", buf[28 + tlen + dlen:32 + tlen + dlen])
    return version, typ, title.decode(), nseq, total, mx


def test_round_trip_and_reference_header_walk(tmp_path):
    rng = np.random.default_rng(0)
    res, offs, lens = random_db(rng, 300)
    base = str(tmp_path / "vol")
    bpx.volume_write(base, res, offs[:-1], lens, title="synthetic volume")
    r2, o2, l2, title = bpx.volume_read(base)
    np.testing.assert_array_equal(l2, lens)
    np.testing.assert_array_equal(o2, offs)
    np.testing.assert_array_equal(r2, res)
    assert title == "synthetic volume"
    version, typ, t, nseq, total, mx = reference_header_walk(open(base + ".pin", "rb").read())
    assert (version, typ, t, nseq, total, mx) == (4, 1, "synthetic volume", 300, int(lens.sum()), int(lens.max()))
    # .psq: a leading separator, then each sequence followed by one
    psq = np.frombuffer(open(base + ".psq", "rb").read(), dtype=np.uint8)
    assert psq[0] == 0 and len(psq) == 1 + int(lens.sum()) + 300


def test_alias_lists_volumes_in_order(tmp_path):
    rng = np.random.default_rng(1)
    parts = [random_db(rng, n) for n in (50, 1, 120)]
    names = []
    for i, (r, o, ln) in enumerate(parts):
        names.append(f"db.{i:02d}")
        bpx.volume_write(str(tmp_path / names[-1]), r, o[:-1], ln)
    (tmp_path / "db.pal").write_text("#\n# Alias file created by a test\n#\nTITLE three volumes\nDBLIST "
                                     + " ".join(names) + "\n")
    vols = bpx.db_volumes(str(tmp_path / "db"))
    assert vols == [str(tmp_path / n) for n in names]
    assert bpx.db_volumes(str(tmp_path / "db.01")) == [str(tmp_path / "db.01")]


def test_malformed_volumes_are_refused(tmp_path):
    rng = np.random.default_rng(2)
    res, offs, lens = random_db(rng, 20)
    base = str(tmp_path / "v")
    bpx.volume_write(base, res, offs[:-1], lens)
    pin = bytearray(open(base + ".pin", "rb").read())
    bad = bytearray(pin)
    bad[7] = 0  # database type 0: nucleotide
    open(str(tmp_path / "n.pin"), "wb").write(bad)
    os.link(base + ".psq", str(tmp_path / "n.psq"))
    with pytest.raises(bpx.BpxError, match="protein"):
        bpx.volume_read(str(tmp_path / "n"))
    open(str(tmp_path / "t.pin"), "wb").write(pin[:len(pin) - 10])
    os.link(base + ".psq", str(tmp_path / "t.psq"))
    with pytest.raises(bpx.BpxError, match="truncated"):
        bpx.volume_read(str(tmp_path / "t"))
    with pytest.raises(bpx.BpxError, match="cannot read"):
        bpx.volume_read(str(tmp_path / "missing"))
    (tmp_path / "e.pal").write_text("TITLE nothing\n")
    with pytest.raises(bpx.BpxError, match="DBLIST"):
        bpx.db_volumes(str(tmp_path / "e"))


def test_hostile_index_headers_are_refused(tmp_path):
    """Header fields that 32-bit arithmetic would wrap: a sequence count of
    2^32 - 1 (n + 1 = 0) and a first sequence offset of 2^32 - 1 (so[0] + 1 =
    0).  Both must be refused with a message -- no endless loop, no
    allocation sized by the count, no read outside the .psq file."""
    rng = np.random.default_rng(3)
    res, offs, lens = random_db(rng, 5)
    base = str(tmp_path / "v")
    bpx.volume_write(base, res, offs[:-1], lens)
    pin = bytearray(open(base + ".pin", "rb").read())
    _, _, title, nseq, _, _ = reference_header_walk(bytes(pin))
    tlen = len(title)
    (dlen,) = struct.unpack(">I", bytes(pin[12 + tlen:16 + tlen]))
    at_n = 16 + tlen + dlen          # sequence count
    at_so = at_n + 4 + 8 + 4 + 4 * (nseq + 1)  # first sequence offset
    huge = bytearray(pin)
    huge[at_n:at_n + 4] = struct.pack(">I", 0xFFFFFFFF)
    open(str(tmp_path / "h.pin"), "wb").write(huge)
    os.link(base + ".psq", str(tmp_path / "h.psq"))
    with pytest.raises(bpx.BpxError, match="truncated"):
        bpx.volume_read(str(tmp_path / "h"))
    wrap = bytearray(pin)
    wrap[at_so:at_so + 4] = struct.pack(">I", 0xFFFFFFFF)
    open(str(tmp_path / "w.pin"), "wb").write(wrap)
    os.link(base + ".psq", str(tmp_path / "w.psq"))
    with pytest.raises(bpx.BpxError, match="outside the file"):
        bpx.volume_read(str(tmp_path / "w"))

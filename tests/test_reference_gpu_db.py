"""The reference's own GPU database files (.gpu / .gpuinfo), read through the
C ABI (host code, no GPU).  The files are produced here by a restatement of
the reference's writer -- the interleave of functions.cpp:364-613 (groups of
blocks * threads sequences in ascending length, group length = its longest
sequence rounded up to 64 (gpu_blastp.c:12-17), residue 4i + j of the m-th
sequence of a group at byte (m + i * G) * 4 + j) and the text records of
functions.cpp:1-46 -- written independently of the library's reader, so a
wrong stride, group base or record order fails the round trip."""
import os

import numpy as np
import pytest

import bpx


def reference_layout(seqs, blocks, threads):
    """CreateGPU_BLASTP_database's byte image for one volume (sequences
    already in ascending length), and its group lengths."""
    G = blocks * threads
    groups = [seqs[g:g + G] for g in range(0, len(seqs), G)]
    glen = []
    for grp in groups:
        L = len(grp[-1])                      # the longest (last) sequence of the group
        if L % 64:
            L += 64 - L % 64
        glen.append(L)
    img = np.zeros(sum(L * G for L in glen), dtype=np.uint8)
    base = 0
    for grp, L in zip(groups, glen):
        for m, s in enumerate(grp):
            for r, a in enumerate(s):
                img[base + (m + (r // 4) * G) * 4 + (r % 4)] = a
        base += L * G
    return img, glen


def write_reference_files(dirpath, name, volumes, blocks, threads, percentage=95, record_dir=True):
    """Writes <name>.gpuinfo and one .gpu file per volume the way
    Write_GPU_information_file_preliminary / Write_GPU_information_file lay
    them out.  volumes: per volume, (all sequences of the volume sorted by
    length, how many go to the .gpu file)."""
    info = [f"#Number of blocksx\n{blocks}\n", f"#Number of threadsx\n{threads}\n",
            f"#Percentage\n{percentage}\n", f"#Number of volumes\n{len(volumes)}\n"]
    break_limit = 0
    for v, (seqs, n_gpu) in enumerate(volumes):
        img, glen = reference_layout(seqs[:n_gpu], blocks, threads)
        leaf = f"{name}.gpu" if len(volumes) == 1 else f"{name}.{v:02d}.gpu"
        img.tofile(os.path.join(dirpath, leaf))
        store_limit = break_limit + n_gpu
        break_limit += len(seqs)
        recorded = os.path.join(dirpath, leaf) if record_dir else os.path.join("/nonexistent/dir", leaf)
        info += [f"#Volume name\n{recorded}\n", f"#Volume size in bytes\n{img.size}\n",
                 f"#Volume sequences\n{len(seqs)}\n", f"#Number of sequences in GPU volume\n{n_gpu}\n",
                 f"#Store limit\n{store_limit}\n", f"#Break limit\n{break_limit}\n",
                 f"#Group number\n{len(glen)}\n", "#Length of each group\n"] + [f"{L}\n" for L in glen]
    path = os.path.join(dirpath, f"{name}.gpuinfo")
    with open(path, "w") as f:
        f.write("".join(info))
    return path


def sorted_seqs(rng, n, lo=1, hi=300):
    lens = np.sort(rng.integers(lo, hi, n))
    return [rng.integers(1, 26, int(L)).astype(np.uint8) for L in lens]


def check_volume(got, seqs):
    res, offs, lens, _ = got
    assert len(lens) == len(seqs)
    for i, s in enumerate(seqs):
        assert int(lens[i]) == len(s)
        np.testing.assert_array_equal(res[int(offs[i]):int(offs[i + 1])], s)


def test_single_volume_with_a_partial_last_group(tmp_path):
    rng = np.random.default_rng(1)
    seqs = sorted_seqs(rng, 250)                     # 250 = 3 groups of 64 + 58
    path = write_reference_files(str(tmp_path), "db", [(seqs, 230)], blocks=4, threads=16)
    assert bpx.gpu_info(path) == (1, 64, 95)
    check_volume(bpx.gpu_volume_read(path, 0), seqs[:230])


def test_lengths_at_stride_and_group_boundaries(tmp_path):
    # lengths 1..3 (inside one 4-residue word), multiples of 4 and of 64, and
    # a group whose longest sequence is an exact multiple of 64
    lens = [1, 2, 3, 4, 5, 63, 64, 64, 65, 127, 128, 129, 256]
    rng = np.random.default_rng(2)
    seqs = [rng.integers(1, 28, L).astype(np.uint8) for L in lens]
    path = write_reference_files(str(tmp_path), "edge", [(seqs, len(seqs))], blocks=2, threads=2)
    check_volume(bpx.gpu_volume_read(path, 0), seqs)


def test_several_volumes_and_a_moved_directory(tmp_path):
    rng = np.random.default_rng(3)
    vols = [(sorted_seqs(rng, 100), 95), (sorted_seqs(rng, 40, 50, 900), 38), (sorted_seqs(rng, 7), 6)]
    # the .gpuinfo records a directory that no longer exists: the reader falls
    # back to the file names next to the .gpuinfo
    path = write_reference_files(str(tmp_path), "nr", vols, blocks=3, threads=5, record_dir=False)
    assert bpx.gpu_info(path)[:2] == (3, 15)
    for v, (seqs, n_gpu) in enumerate(vols):
        check_volume(bpx.gpu_volume_read(path, v), seqs[:n_gpu])


def test_malformed_files_fail_with_a_message(tmp_path):
    rng = np.random.default_rng(4)
    seqs = sorted_seqs(rng, 20)
    path = write_reference_files(str(tmp_path), "bad", [(seqs, 20)], blocks=2, threads=4)
    with pytest.raises(bpx.BpxError, match="out of range"):
        bpx.gpu_volume_read(path, 1)
    with pytest.raises(bpx.BpxError, match="cannot read"):
        bpx.gpu_volume_read(str(tmp_path / "missing.gpuinfo"), 0)
    # truncate the .gpu file: its size no longer matches the record
    gpu = str(tmp_path / "bad.gpu")
    data = np.fromfile(gpu, dtype=np.uint8)
    data[:-64].tofile(gpu)
    with pytest.raises(bpx.BpxError, match="size differs"):
        bpx.gpu_volume_read(path, 0)
    # a wrong group count in the record
    text = open(path).read().replace("#Group number\n3\n", "#Group number\n2\n", 1)
    open(path, "w").write(text)
    with pytest.raises(bpx.BpxError):
        bpx.gpu_volume_read(path, 0)

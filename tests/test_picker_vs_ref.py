"""The served form's compaction picker (luda_amd/served.py, Lsm.pick) against
the reference's own VersionSet::PickCompaction (db/version_set.cc:1270-1464,
with Finalize's scores, AddBoundaryInputs, SetupOtherInputs' expansion,
grandparents and IsTrivialMove), run from oracle/_ref on the same random
trees: the same level, input files in the same order, grandparents, trivial
moves and compact pointer.  Host code only."""
import ctypes as C
import random
import struct

import pytest

import oracle
from luda_amd import ffi, served

pytestmark = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference tree)")

MIB = 1 << 20


def ref_pick(tmpdir, tree_files, pointers, max_file_size):
    """tree_files: [(level, number, size, smallest, largest)]; pointers:
    [(level, internal key)]."""
    lib = oracle.load_ref()
    f = lib.ref_pick
    f.restype = C.c_int
    keep = []

    def sl(b):
        buf = C.create_string_buffer(b, len(b))
        keep.append(buf)
        return ffi.luda_slice(C.cast(buf, C.c_void_p), len(b))
    nf = (ffi.luda_edit_file * max(1, len(tree_files)))()
    for i, (lv, num, size, lo, hi) in enumerate(tree_files):
        nf[i].level, nf[i].number, nf[i].file_size = lv, num, size
        nf[i].smallest, nf[i].largest = sl(lo), sl(hi)
    cps = (ffi.luda_edit_pointer * max(1, len(pointers)))()
    for i, (lv, k) in enumerate(pointers):
        cps[i].level, cps[i].key = lv, sl(k)
    e = ffi.luda_edit()
    e.new_files, e.num_new_files = nf, len(tree_files)
    e.compact_pointers, e.num_compact_pointers = cps, len(pointers)
    cap = len(tree_files) + 1
    in0, in1, gp = (C.c_uint64 * cap)(), (C.c_uint64 * cap)(), (C.c_uint64 * cap)()
    n0, n1, ngp = C.c_uint32(), C.c_uint32(), C.c_uint32()
    level, trivial = C.c_int(), C.c_int()
    edit = C.create_string_buffer(4096)
    elen = C.c_uint64()
    rc = f(str(tmpdir).encode(), C.byref(e), C.c_uint64(max_file_size), C.byref(level), in0, C.byref(n0), in1,
           C.byref(n1), gp, C.byref(ngp), C.byref(trivial), edit, C.c_uint64(4096), C.byref(elen))
    assert rc == 0
    return (level.value, list(in0[:n0.value]), list(in1[:n1.value]), list(gp[:ngp.value]), bool(trivial.value),
            edit.raw[:elen.value])


def ikey(user, seq, t=1):
    return user + struct.pack("<Q", (seq << 8) | t)


def random_tree(rng):
    """Level 0: overlapping files; levels 1-6: sorted files, some sharing a
    boundary user key with their neighbour (older versions in the next
    file); sizes scaled to each level's target so several scores pass 1."""
    files = []
    num = [10]

    def add(level, lo, hi, size):
        num[0] += rng.randint(1, 3)
        files.append((level, num[0], size, lo, hi))
    seq = 10 ** 6
    for _ in range(rng.choice([0, 1, 3, 4, 5, 8])):
        a = rng.randrange(0, 9000)
        b = a + rng.randrange(0, 3000)
        add(0, ikey(b"k%05d" % a, seq), ikey(b"k%05d" % b, seq - 1), rng.randint(1, 4) * MIB)
        seq -= 2
    for level in range(1, 7):
        if rng.random() < 0.25:
            continue
        target = 10.0 * MIB * 10 ** (level - 1)
        nfiles = rng.randint(1, 14)
        cuts = sorted(rng.sample(range(1, 10000), 2 * nfiles))
        share = False
        for i in range(nfiles):
            lo_u, hi_u = cuts[2 * i], cuts[2 * i + 1]
            lo = ikey(b"k%05d" % lo_u, seq)
            if share:  # smallest shares the previous file's largest user key
                lo = ikey(files[-1][4][:-8], seq)
            seq -= 1
            share = rng.random() < 0.3 and i + 1 < nfiles
            hi = ikey(b"k%05d" % hi_u, seq)
            seq -= 1
            size = int(target * rng.uniform(0.5, 2.0) / nfiles)
            add(level, lo, hi, max(1, size))
    pointers = []
    for level in range(6):
        if rng.random() < 0.4:
            pointers.append((level, ikey(b"k%05d" % rng.randrange(10000), rng.randrange(1, seq))))
    return files, pointers


def python_pick(files, pointers, max_file_size):
    lsm = served.Lsm(max_file_size=max_file_size, level1_bytes=10 * MIB)
    where = {}
    for level, number, size, lo, hi in files:
        lsm.add_file(level, served.File(number, size, lo, hi, 1, None))
        where[number] = level
    for level, k in pointers:
        lsm.compact_pointer[level] = k
    got = lsm.pick()
    if got is None:
        return -1, [], [], [], False, None
    if isinstance(got, served.File):
        level = where[got.number]
        # the moved file's level-2 overlap is its grandparents
        return level, [got.number], [], None, True, lsm.compact_pointer[level]
    return (got.level, [f.number for f in got.inputs[0]], [f.number for f in got.inputs[1]],
            [f.number for f in got.grandparents], False, lsm.compact_pointer[got.level])


def encode_pointer(level, key):
    lib = ffi.engine()
    b = C.create_string_buffer(key, len(key))
    cp = (ffi.luda_edit_pointer * 1)()
    cp[0].level, cp[0].key = level, ffi.luda_slice(C.cast(b, C.c_void_p), len(key))
    e = ffi.luda_edit()
    e.compact_pointers, e.num_compact_pointers = cp, 1
    out = C.create_string_buffer(4096)
    n = C.c_uint64()
    assert lib.luda_encode_edit(C.byref(e), out, 4096, C.byref(n)) == 0
    return out.raw[:n.value]


@pytest.mark.parametrize("block", range(6))
def test_picks_equal_reference(tmp_path, block):
    counts = {"none": 0, "trivial": 0, "job": 0, "expanded": 0}
    for seed in range(block * 60, (block + 1) * 60):
        rng = random.Random(seed)
        mfs = rng.choice([2 * MIB, 32 * MIB, 256 * MIB])
        files, pointers = random_tree(rng)
        want = ref_pick(tmp_path / f"db{seed}", files, pointers, mfs)
        got = python_pick(files, pointers, mfs)
        assert got[0] == want[0], seed
        if want[0] < 0:
            counts["none"] += 1
            continue
        assert got[4] == want[4], seed
        assert got[1] == want[1] and got[2] == want[2], seed
        if not got[4]:
            assert got[3] == want[3], seed
        assert encode_pointer(got[0], got[5]) == want[5], seed
        counts["trivial" if got[4] else "job"] += 1
    assert counts["job"] > 10 and counts["trivial"] > 0, counts

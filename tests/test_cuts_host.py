"""The tiled output block cuts (luda_amd/csrc/cuts.h, used by k_cuts.hip),
built for the host with address and undefined-behaviour sanitizers, against
a one-entry-at-a-time restatement of TableBuilder's block flush
(table/table_builder.cc:119-123) and DoCompactionWork's cut before a
ShouldStopBefore entry (db/db_impl.cc:1105-1111): tests/cuts_harness.cc.

Shapes: entries of ~120 bytes (the benchmark's), of 20-300 bytes, mixed with
large values, all large (one entry per block), and of 4-16 bytes (more entry
points than a tile takes: the engine falls back to the orbit kernels, which
the harness reports); restart intervals 1-64; block sizes from 1 byte to
64 KiB; streams that end on, just past and just before tile edges; forced
table starts from none to one every few entries, one on a tile's first
entry.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "luda_amd", "csrc")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    exe = str(tmp_path_factory.mktemp("cuts") / "cuts_harness")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-I", CSRC, os.path.join(HERE, "cuts_harness.cc"), "-o", exe],
                   check=True)
    return exe


def run(exe, seed, m, R, bs, shape, forced_every):
    p = subprocess.run([exe, str(seed), str(m), str(R), str(bs), shape, str(forced_every)], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.startswith("ok "), p.stdout
    return int(p.stdout.split("fallback=")[1])


CASES = [
    (1, 14000, 16, 4096, "fixed", 0, 0),
    (2, 14000, 16, 4096, "fixed", 5000, 0),
    (3, 30000, 1, 4096, "varied", 300, 0),
    (4, 30000, 64, 4096, "varied", 0, 0),
    (6, 9000, 16, 4096, "huge", 0, 0),
    (7, 20000, 2, 4096, "mixed", 40, 0),
    (8, 1, 16, 4096, "fixed", 0, 0),
    (9, 4096, 16, 4096, "fixed", 0, 0),
    (10, 4097, 16, 4096, "fixed", 1, 0),
    (11, 5121, 16, 4096, "varied", 0, 0),
    (16, 8191, 16, 4096, "varied", 3, 0),
    (12, 9000, 1, 1, "varied", 0, 0),
    (13, 9000, 4, 64, "varied", 7, 0),
    (15, 50000, 16, 4096, "varied", 0, 0),
    (17, 40000, 16, 65536, "varied", 0, 1),  # ~400 entries per block
    (18, 40000, 16, 65536, "huge", 0, 0),
    (5, 20000, 16, 4096, "small", 0, 1),
    (14, 20000, 16, 16384, "small", 0, 1),
]


@pytest.mark.parametrize("seed,m,R,bs,shape,forced,fallback", CASES)
def test_tiled_cuts_match_table_builder(harness, seed, m, R, bs, shape, forced, fallback):
    assert run(harness, seed, m, R, bs, shape, forced) == fallback


def test_tiled_cuts_random_sweep(harness):
    tiled = 0
    for seed in range(100, 160):
        for shape in ("fixed", "varied", "mixed", "huge"):
            R = 1 if seed % 5 == 0 else (64 if seed % 3 == 0 else 16)
            bs = 256 if seed % 7 == 0 else (65536 if seed % 11 == 0 else 4096)
            m = (seed * 7919) % 40000 + 1
            f = 0 if seed % 4 == 0 else seed % 200 + 2
            tiled += run(harness, seed, m, R, bs, shape, f) == 0
    assert tiled > 200  # most shapes take the tiled path

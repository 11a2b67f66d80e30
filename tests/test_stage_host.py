"""The block writer's value staging (luda_amd/csrc/stage_values.h, used by
k_write_blocks) built for the host with address and undefined-behaviour
sanitizers: values stored as aligned dwords with junk at their ends (and long
values by 16-byte slots), then headers, keys and restart words, must leave
exactly the block BlockBuilder lays out (table/block_builder.cc:62-114), with
every stage write inside the stage and every value read within the 15 bytes
past the value that its source allows: tests/stage_harness.cc.

Shapes: values of 0-128 bytes (each lane its own value, loaded 2 or 4 chunks
at a time), of 129-1128 bytes (spread by slots), mixed, and of 0-5 bytes with
headers of 3 bytes and empty key deltas (the least room for junk); restart
intervals 1-40; blocks of one entry to three rounds of 64; every offset of the
block in its 16-byte stage slot.
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
    exe = str(tmp_path_factory.mktemp("stage") / "stage_harness")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-I", CSRC, os.path.join(HERE, "stage_harness.cc"), "-o", exe],
                   check=True)
    return exe


@pytest.mark.parametrize("seed", range(8))
def test_stage_matches_block_builder(harness, seed):
    p = subprocess.run([harness, str(seed), "3000"], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.startswith("ok "), p.stdout


def test_planted_junk_is_caught(harness, tmp_path):
    """A harness whose header phase skips one byte must fail: the junk the
    values leave is only gone once the headers overwrite it."""
    src = open(os.path.join(HERE, "stage_harness.cc")).read()
    old = "for (uint32_t t = 0; t < es[j].head.size(); t++) w8(off[j] + t, es[j].head[t]);"
    assert old in src
    bad = src.replace(old, "for (uint32_t t = 0; t < es[j].head.size(); t++) if (t + 1 != es[j].head.size()) "
                           "w8(off[j] + t, es[j].head[t]);")
    f = tmp_path / "bad.cc"
    f.write_text(bad)
    exe = str(tmp_path / "bad")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", CSRC, "-I", HERE, str(f), "-o", exe], check=True)
    p = subprocess.run([exe, "1", "300"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 1 and "FAIL" in p.stdout, p.stdout

"""The served form's host side (luda_amd/served.py) on the CPU, with the oracle
doing the table work: the tree it keeps is a valid LevelDB tree and holds
exactly the newest put of every key, run one job at a time and with queued
jobs finishing out of order."""
import pytest

from luda_amd import served
from served_backends import OracleBackend, expected_contents, small_options, tree_contents


def stream(puts=40000, key_space=30000, epf=1500, seed=7):
    return list(served.write_stream(puts, seed, key_space=key_space, entries_per_flush=epf))


def test_write_stream_memtables_sorted():
    mems = stream(puts=5000, epf=1200)
    assert [m.n for m in mems] == [1200, 1200, 1200, 1200, 200]
    seq = 1
    for m in mems:
        assert m.first_seq == seq
        seq = m.last_seq + 1
        keys = [bytes(k) for k in m.keys]
        assert all(len(k) == 24 and k[:16].isdigit() for k in keys)
        assert keys == sorted(keys, key=served.ikey_order)
        tags = sorted(int.from_bytes(k[16:], "little") for k in keys)
        assert tags == [(s << 8) | 1 for s in range(m.first_seq, m.last_seq + 1)]


def test_max_bytes_for_level():
    assert served.max_bytes_for_level(1) == 10 * 1048576
    assert served.max_bytes_for_level(4) == 10000 * 1048576


@pytest.mark.parametrize("mode", ["plain", "served"])
def test_tree_holds_newest_puts(mode):
    mems = stream()
    opts = small_options()
    be = OracleBackend(opts)
    lsm = served.Lsm(max_file_size=opts.max_file_size, level1_bytes=48 << 10)
    seen = []

    def on_job(spec, files):
        # every compaction writes its outputs sorted and disjoint
        for a, b in zip(files, files[1:]):
            assert served.ikey_order(a.largest) < served.ikey_order(b.smallest)
        seen.append((spec.level, len(spec.inputs[0]), len(spec.inputs[1])))

    run = served.run_plain if mode == "plain" else served.run_served
    lsm, st = run(be, mems, 0.0, lsm=lsm, on_job=on_job)
    lsm.check()
    summ = st.summary()
    assert summ["flushes"] == len(mems)
    assert summ["compactions"] == len(seen) > 0
    assert {level for level, _, _ in seen} >= {0, 1}
    assert lsm.levels[3] and lsm.moves > 0, "the stream never reached level 3"
    assert not lsm.needs()
    assert tree_contents(lsm, be.table_bytes) == expected_contents(mems)


def test_picker_rules():
    """Level-0 inputs are widened to every overlapping level-0 file, level+1
    inputs are the overlapping files, and a lone file with nothing below it
    moves down."""
    lsm = served.Lsm(max_file_size=1 << 20, level1_bytes=1 << 20)

    def f(lo, hi, size=1000, seq=1):
        tag = ((seq << 8) | 1).to_bytes(8, "little")
        return served.File(lsm.new_number(), size, lo.encode() + tag, hi.encode() + tag, 1, None)
    a, b, c, d = f("b", "d"), f("c", "f"), f("e", "g"), f("x", "z")
    for x in (a, b, c, d):
        lsm.add_file(0, x)
    l1 = [f("a", "c"), f("d", "e"), f("h", "k")]
    for x in l1:
        lsm.add_file(1, x)
    spec = lsm.pick()
    assert spec.level == 0
    assert set(spec.inputs[0]) == {a, b, c}
    assert set(spec.inputs[1]) == {l1[0], l1[1]}
    assert lsm.pick() is None  # levels 0 and 1 are in use
    lsm.install(spec, [f("a", "g")])
    assert [x for x in lsm.levels[0]] == [d]
    lsm.check()
    # a single level-1 file over the limit with no level-2 overlap moves down
    lsm2 = served.Lsm(max_file_size=1 << 20, level1_bytes=1000)
    lsm = lsm2
    g = f("m", "n", size=5000)
    lsm2.add_file(1, g)
    assert lsm2.pick() is g
    assert lsm2.levels[2] == [g] and lsm2.moves == 1


def test_deeper_rows_match_the_job():
    """A job carries, as luda_level_file rows, exactly the files of levels >=
    level + 2 that overlap its key range (IsBaseLevelForKey is only asked
    about the job's keys), the same rows _job_struct builds from the files,
    and keeps them when a deeper level changes after the pick."""
    import ctypes as C
    import numpy as np
    lsm = served.Lsm(max_file_size=1 << 20, level1_bytes=1000)

    def f(lo, hi, size=600, seq=1):
        tag = ((seq << 8) | 1).to_bytes(8, "little")
        return served.File(lsm.new_number(), size, lo.encode() + tag, hi.encode() + tag, 1, 0)
    lsm.add_file(1, f("c", "e"))
    lsm.add_file(1, f("f", "h"))
    lsm.add_file(2, f("d", "g"))
    for lvl, lo, hi in [(3, "a0", "b9"), (3, "c5", "d5"), (3, "g0", "g9"), (4, "e0", "e9"), (4, "x0", "x9"),
                        (5, "a", "z"), (6, "h5", "h9")]:
        lsm.add_file(lvl, f(lo, hi))
    spec = lsm.pick()
    # c .. e takes the level-2 file d .. g; growing the level-1 inputs to
    # f .. h takes no other level-2 file, so SetupOtherInputs takes both
    # (db/version_set.cc:1420-1450)
    assert spec.level == 1 and [x.smallest[:-8] for x in spec.inputs[0]] == [b"c", b"f"]
    # the job spans user keys c .. h; "h5" sorts after "h", so the level-6
    # file h5 .. h9 is outside
    got = [(d, x.smallest[:-8].decode()) for d, x in spec.deeper]
    assert got == [(3, "c5"), (3, "g0"), (4, "e0"), (5, "a")]
    opts = small_options()

    def rows(job):
        return bytes(C.string_at(C.cast(job.deeper, C.c_void_p), job.num_deeper * C.sizeof(served.ffi.luda_level_file)))
    with_rows, _k1 = served._job_struct(spec, opts)
    rebuilt = served.JobSpec(spec.level, spec.inputs, spec.grandparents, spec.deeper_files, spec.smallest_snapshot)
    from_files, _k2 = served._job_struct(rebuilt, opts)
    assert rows(with_rows) == rows(from_files) and with_rows.num_deeper == 4
    lsm.add_file(4, f("ea", "eb"))  # a deeper level changes after the pick
    again, _k3 = served._job_struct(spec, opts)
    assert rows(again) == rows(with_rows)
    assert np.frombuffer(rows(with_rows), dtype=served._LEVEL_FILE)["level"].tolist() == [3, 3, 4, 5]


@pytest.mark.parametrize("mode", ["plain", "served"])
def test_a_failed_job_stops_the_run(mode):
    """A compaction the backend rejects ends the run with its error instead
    of leaving the tree waiting on a job that never completes."""
    mems = stream(puts=12000, key_space=9000)
    opts = small_options()

    class Failing(OracleBackend):
        def compact(self, lsm, spec):
            raise RuntimeError("compaction rejected")
    be = Failing(opts)
    lsm = served.Lsm(max_file_size=opts.max_file_size, level1_bytes=48 << 10)
    run = served.run_plain if mode == "plain" else served.run_served
    with pytest.raises(RuntimeError, match="compaction rejected"):
        run(be, mems, 0.0, lsm=lsm)

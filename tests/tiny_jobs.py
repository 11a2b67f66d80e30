"""Seeded tiny compaction jobs for brute-force differential tests: a handful
of short keys, many versions, deletions, snapshots, grandparent cuts and
deeper-level ranges, with tiny block and file sizes so that every branch of
the compaction loop (db/db_impl.cc:1090-1184) and of the table builder is
reached."""
import ctypes as C
import random
import struct

from luda_amd import ffi

USER_KEYS = [b"a", b"ab", b"abc", b"abd", b"b", b"ba", b"b\xff", b"c", b"c\x00", b"cz", b"d\xff\xff", b"\xff",
             b"\xff\xff", b"key0001", b"key0002", b"key0010", b"longer-user-key-000", b"longer-user-key-001"]


def ikey(user, seq, t):
    return user + struct.pack("<Q", (seq << 8) | t)


def _run(rng, seq, nkeys, max_versions, del_p):
    ents = []
    keys = rng.sample(USER_KEYS, min(nkeys, len(USER_KEYS)))
    for u in keys:
        for _ in range(rng.randint(1, max_versions)):
            seq[0] += 1
            t = 0 if rng.random() < del_p else 1
            ents.append((u, seq[0], t))
    rng.shuffle(ents)
    # give sequence numbers in insertion order
    out = []
    for u, _, t in ents:
        seq[0] += 1
        out.append((u, seq[0], t))
    out.sort(key=lambda e: (e[0], -e[1]))
    return out


def _entries(run, rng):
    return [(ikey(u, s, t), b"" if t == 0 else bytes(rng.randrange(256) for _ in range(rng.choice([0, 1, 5, 40, 130]))))
            for u, s, t in run]


def _split(entries, rng, k):
    if k <= 1 or len(entries) < 2:
        return [entries]
    cuts = sorted(rng.sample(range(1, len(entries)), min(k - 1, len(entries) - 1)))
    parts, prev = [], 0
    for c in cuts + [len(entries)]:
        parts.append(entries[prev:c])
        prev = c
    return parts


class TinyJob:
    def __init__(self, seed):
        rng = random.Random(seed)
        self.seed = seed
        self.level = rng.choice([0, 0, 1, 2, 4])
        self.opts = ffi.default_options(block_size=rng.choice([32, 64, 128, 300, 4096]),
                                        block_restart_interval=rng.choice([1, 2, 4, 16]),
                                        bloom_bits_per_key=rng.choice([0, 10, 10, 3]),
                                        max_file_size=rng.choice([150, 400, 1000, 1 << 20]))
        seq = [0]
        older = _entries(_run(rng, seq, rng.randint(0, 8), 3, 0.3), rng)
        self.files = ([], [])
        self.files[1].extend(p for p in _split(older, rng, rng.randint(1, 3)) if p)
        if self.level == 0:
            for _ in range(rng.randint(1, 4)):
                r = _entries(_run(rng, seq, rng.randint(1, 6), 3, 0.3), rng)
                self.files[0].append(r)
        else:
            r = _entries(_run(rng, seq, rng.randint(1, 8), 3, 0.3), rng)
            self.files[0].extend(p for p in _split(r, rng, rng.randint(1, 2)) if p)
        self.snapshot = rng.randint(0, seq[0] + 1)
        # grandparents: sorted internal keys with sizes near the cut budget
        gk = sorted(rng.sample(USER_KEYS, rng.randint(0, 6)))
        self.grandparents = [(ikey(u, rng.randint(1, 50), 1), rng.choice([100, 800, 1500, 3000])) for u in gk]
        self.deeper = []
        for lvl in range(self.level + 2, 7):
            if rng.random() < 0.5:
                continue
            ks = sorted(rng.sample(USER_KEYS, rng.choice([0, 2, 4])))
            for i in range(0, len(ks) - 1, 2):
                self.deeper.append((lvl, ks[i], ks[i + 1]))

    def make_job(self, build, device_inputs=None):
        """build(entries, opts) -> [table bytes]; returns (luda_job, keepalive)."""
        keep = []
        job = ffi.luda_job()
        job.level = self.level
        job.smallest_snapshot = self.snapshot
        job.options = self.opts
        self.tables = ([], [])
        one = ffi.default_options()
        C.pointer(one)[0] = self.opts
        one.max_file_size = ffi.U64_MAX
        for which in (0, 1):
            for ents in self.files[which]:
                self.tables[which].extend(build(ents, one))
        for which in (0, 1):
            ts = self.tables[which]
            arr = (ffi.luda_file * max(1, len(ts)))()
            for i, t in enumerate(ts):
                if device_inputs is not None:
                    ptr = device_inputs(t)
                else:
                    b = C.create_string_buffer(t, len(t))
                    keep.append(b)
                    ptr = C.cast(b, C.c_void_p)
                arr[i].data = ptr
                arr[i].size = len(t)
                arr[i].number = i + 1
            keep.append(arr)
            job.inputs[which] = C.cast(arr, C.POINTER(ffi.luda_file))
            job.num_inputs[which] = len(ts)
        if device_inputs is not None:
            job.flags |= ffi.LUDA_JOB_INPUTS_ON_DEVICE
        gps = (ffi.luda_grandparent * max(1, len(self.grandparents)))()
        for i, (k, sz) in enumerate(self.grandparents):
            b = C.create_string_buffer(k, len(k))
            keep.append(b)
            gps[i].largest.data = C.cast(b, C.c_void_p)
            gps[i].largest.size = len(k)
            gps[i].file_size = sz
        keep.append(gps)
        job.grandparents = C.cast(gps, C.POINTER(ffi.luda_grandparent))
        job.num_grandparents = len(self.grandparents)
        dp = (ffi.luda_level_file * max(1, len(self.deeper)))()
        for i, (lvl, s, l) in enumerate(self.deeper):
            bs = C.create_string_buffer(s, len(s))
            bl = C.create_string_buffer(l, len(l))
            keep += [bs, bl]
            dp[i].smallest_user_key.data = C.cast(bs, C.c_void_p)
            dp[i].smallest_user_key.size = len(s)
            dp[i].largest_user_key.data = C.cast(bl, C.c_void_p)
            dp[i].largest_user_key.size = len(l)
            dp[i].level = lvl
        keep.append(dp)
        job.deeper = C.cast(dp, C.POINTER(ffi.luda_level_file))
        job.num_deeper = len(self.deeper)
        return job, keep


def many_grandparents_job(ngp, seed=5):
    """A level-1 job over 6000 user keys whose grandparents (ngp of them, small
    files) interleave the keys, so ShouldStopBefore fires every few of them
    and the crossings outnumber one LDS pass of the cut walk."""
    rng = random.Random(seed)
    users = [b"k%06d" % i for i in range(6000)]
    seq = 0
    a, b = [], []
    for u in users:
        seq += 1
        (a if rng.random() < 0.3 else b).append((ikey(u, seq, 1), bytes(rng.randrange(256) for _ in range(20))))
    tj = TinyJob(0)
    tj.level = 1
    tj.opts = ffi.default_options(max_file_size=4096)
    tj.files = ([a], [b])
    tj.snapshot = seq
    picks = sorted(rng.sample(range(6000), ngp))
    tj.grandparents = [(ikey(users[i] + b"~", 7, 1), rng.choice([900, 2500, 6000])) for i in picks]
    tj.deeper = []
    return tj


def _tj_shell(seed, level, opts):
    tj = TinyJob(0)
    tj.seed = seed
    tj.level = level
    tj.opts = opts
    return tj


def many_children_job(nl0, seed):
    """A level-0 job with nl0 overlapping L0 files (each its own merge child,
    db/version_set.cc:1245-1252) over a shared key set, plus 1-3 L1 files:
    ceil(log2(nl0 + 1)) merge rounds, versions spread over many children."""
    rng = random.Random(seed)
    opts = ffi.default_options(block_size=rng.choice([128, 512, 4096]),
                               block_restart_interval=rng.choice([1, 4, 16]),
                               bloom_bits_per_key=rng.choice([0, 10]),
                               max_file_size=rng.choice([3000, 20000, 1 << 20]))
    tj = _tj_shell(seed, 0, opts)
    users = [b"u%05d" % i for i in range(600)]
    seq = 0
    l1 = []
    for u in sorted(rng.sample(users, 200)):
        seq += 1
        l1.append((u, seq, 1))
    l1e = _entries(l1, rng)
    files0 = []
    for _ in range(nl0):
        run = []
        for u in rng.sample(users, rng.randint(10, 60)):
            for _ in range(rng.randint(1, 2)):
                seq += 1
                run.append((u, seq, 0 if rng.random() < 0.2 else 1))
        run.sort(key=lambda e: (e[0], -e[1]))
        files0.append(_entries(run, rng))
    tj.files = (files0, [p for p in _split(l1e, rng, rng.randint(1, 3)) if p])
    tj.snapshot = rng.randint(seq // 2, seq + 1)
    gk = sorted(rng.sample(users, rng.randint(0, 8)))
    tj.grandparents = [(ikey(u, rng.randint(1, 50), 1), rng.choice([800, 3000, 9000])) for u in gk]
    tj.deeper = []
    if rng.random() < 0.5:
        ks = sorted(rng.sample(users, 6))
        tj.deeper = [(2, ks[0], ks[1]), (2, ks[2], ks[3]), (4, ks[4], ks[5])]
    return tj


def wide_key_job(seed, lo=65, hi=256, nl0=5):
    """User keys of lo..hi bytes (key planes wider than 8 words: the grouped
    decoder and the untiled merge) sharing long prefixes, over nl0 L0 files and
    two L1 files."""
    rng = random.Random(seed)
    opts = ffi.default_options(block_size=rng.choice([512, 4096]),
                               block_restart_interval=rng.choice([2, 16]),
                               max_file_size=rng.choice([6000, 1 << 20]))
    tj = _tj_shell(seed, 0, opts)
    stems = [bytes(rng.randrange(256) for _ in range(rng.randint(lo - 8, lo))) for _ in range(4)]
    users = set()
    while len(users) < 120:
        s = rng.choice(stems)
        L = rng.randint(lo, hi)
        users.add((s + bytes(rng.randrange(256) for _ in range(L)))[:L])
    users = sorted(users)
    seq = 0
    l1 = []
    for u in users[::2]:
        seq += 1
        l1.append((u, seq, 1))
    files0 = []
    for _ in range(nl0):
        run = []
        for u in rng.sample(users, 30):
            seq += 1
            run.append((u, seq, 0 if rng.random() < 0.2 else 1))
        run.sort(key=lambda e: (e[0], -e[1]))
        files0.append(_entries(run, rng))
    l1e = _entries(l1, rng)
    tj.files = (files0, [l1e[:len(l1e) // 2], l1e[len(l1e) // 2:]])
    tj.snapshot = seq
    tj.grandparents = []
    tj.deeper = []
    return tj


MALFORMED_TYPES = (2, 3, 0x7F, 0x80, 0xFF)


def malformed_job(seed):
    """A TinyJob in which about a fifth of the entries carry a tag type above
    kTypeValue: ParseInternalKey fails for them, so DoCompactionWork keeps
    them and forgets the current user key (db/db_impl.cc:1113-1119)."""
    tj = TinyJob(seed)
    rng = random.Random(seed ^ 0x5A5A)

    def spoil(ents):
        out = []
        for k, v in ents:
            if rng.random() < 0.2:
                tag = struct.unpack("<Q", k[-8:])[0]
                k = k[:-8] + struct.pack("<Q", (tag & ~0xFF) | rng.choice(MALFORMED_TYPES))
            out.append((k, v))
        # the tag's type byte takes part in the order: re-sort by user key,
        # then tag descending (db/dbformat.cc:51-67)
        out.sort(key=lambda e: (e[0][:-8], -struct.unpack("<Q", e[0][-8:])[0]))
        return out

    tj.files = ([spoil(e) for e in tj.files[0]], [spoil(e) for e in tj.files[1]])
    # files of one level must stay disjoint and sorted across the level
    return tj


def interval_job(seed, interval, block_size=8192):
    """A level-1 job whose tables (inputs and outputs) use restart interval
    `interval`, with hundreds of 10-20 byte entries per block, so restart
    groups reach the interval."""
    rng = random.Random(seed)
    opts = ffi.default_options(block_size=block_size, block_restart_interval=interval,
                               max_file_size=rng.choice([30000, 1 << 20]))
    tj = _tj_shell(seed, 1, opts)
    users = [b"k%06d" % i for i in range(4000)]
    seq = 0
    older, newer = [], []
    for u in users:
        if rng.random() < 0.7:
            seq += 1
            older.append((u, seq, 1))
    for u in sorted(rng.sample(users, 1500)):
        seq += 1
        newer.append((u, seq, 0 if rng.random() < 0.1 else 1))

    def ents(run):
        return [(ikey(u, s, t), b"" if t == 0 else bytes(rng.randrange(256) for _ in range(rng.randint(0, 3))))
                for u, s, t in run]

    oe = ents(older)
    tj.files = ([ents(newer)], [oe[:len(oe) // 2], oe[len(oe) // 2:]])
    tj.snapshot = seq
    tj.grandparents = []
    tj.deeper = []
    return tj

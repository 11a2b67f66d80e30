"""Known-answer checks of the oracle (oracle/lsm_oracle.c) against the
vectors the reference's own tests hold, and against closed forms."""
import ctypes as C
import struct

import pytest

import oracle

O = oracle.load()


def crc(b):
    return O.or_crc32c_value(bytes(b), len(b))


def test_crc32c_rfc3720_vectors():
    # util/crc32c_test.cc:13-41 (RFC 3720 section B.4)
    assert crc(bytes(32)) == 0x8a9136aa
    assert crc(b"\xff" * 32) == 0x62a8ab43
    assert crc(bytes(range(32))) == 0x46dd794e
    assert crc(bytes(31 - i for i in range(32))) == 0x113fdb5c
    data = bytes([0x01, 0xc0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0x14, 0, 0, 0, 0, 0, 0x04, 0, 0, 0, 0,
                  0x14, 0, 0, 0, 0x18, 0x28, 0, 0, 0, 0, 0, 0, 0, 0x02, 0, 0, 0, 0, 0, 0, 0])
    assert crc(data) == 0xd9963a56


def test_crc32c_extend_and_mask():
    # util/crc32c_test.cc:43-54
    assert crc(b"a") != crc(b"foo")
    assert crc(b"hello world") == O.or_crc32c_extend(crc(b"hello "), b"world", 5)
    c = crc(b"foo")
    m = O.or_crc32c_mask(c)
    assert m != c and O.or_crc32c_mask(m) != c
    assert O.or_crc32c_unmask(m) == c
    assert O.or_crc32c_unmask(O.or_crc32c_unmask(O.or_crc32c_mask(m))) == c
    # closed form of util/crc32c.h:29-32: rotate right 15, add 0xa282ead8
    assert m == ((((c >> 15) | (c << 17)) & 0xffffffff) + 0xa282ead8) & 0xffffffff


def test_crc32c_matches_bitwise_definition():
    # Castagnoli CRC, reflected 0x82F63B78, init/final xor ~0, bit by bit.
    def ref(b):
        r = 0xffffffff
        for x in b:
            r ^= x
            for _ in range(8):
                r = (r >> 1) ^ (0x82F63B78 if r & 1 else 0)
        return r ^ 0xffffffff
    for n in (0, 1, 7, 64, 4097):
        b = bytes((i * 131 + 7) & 0xff for i in range(n))
        assert crc(b) == ref(b)


def test_hash_known_values():
    # util/hash_test.cc:12-40
    h = lambda b, s: O.or_hash(bytes(b), len(b), s)
    assert h(b"", 0xbc9f1d34) == 0xbc9f1d34
    assert h(b"\x62", 0xbc9f1d34) == 0xef1345c4
    assert h(b"\xc3\x97", 0xbc9f1d34) == 0x5b663814
    assert h(b"\xe2\x99\xa5", 0xbc9f1d34) == 0x323c078f
    assert h(b"\xe1\x80\xb9\x32", 0xbc9f1d34) == 0xed21633a
    data5 = bytes([0x01, 0xc0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0x14, 0, 0, 0, 0, 0, 0x04, 0, 0, 0, 0,
                   0x14, 0, 0, 0, 0x18, 0x28, 0, 0, 0, 0, 0, 0, 0, 0x02, 0, 0, 0, 0, 0, 0, 0])
    assert h(data5, 0x12345678) == 0xf333dabb


def test_varint_encoding():
    # util/coding_test.cc:78-115: LEB128 round trips and lengths
    buf = C.create_string_buffer(10)
    for v in [0, 1, 127, 128, 255, 300, (1 << 14) - 1, 1 << 14, (1 << 32) - 1, 1 << 32, (1 << 63), (1 << 64) - 1]:
        n = O.or_encode_varint64(buf, v)
        assert n == O.or_varint_length(v)
        got, shift = 0, 0
        for i in range(n):
            b = buf.raw[i]
            got |= (b & 0x7f) << shift
            shift += 7
            assert (b & 0x80 != 0) == (i < n - 1)
        assert got == v
    assert O.or_varint_length((1 << 64) - 1) == 10


def ikey(user, seq, t):
    return user + struct.pack("<Q", (seq << 8) | t)


MAXSEQ = (1 << 56) - 1


def shorten(a, b):
    out = C.create_string_buffer(len(a) + 16)
    n = O.or_shortest_separator(a, len(a), b, len(b), out)
    return out.raw[:n]


def successor(a):
    out = C.create_string_buffer(len(a) + 16)
    n = O.or_short_successor(a, len(a), out)
    return out.raw[:n]


def test_internal_key_short_separator():
    # db/dbformat_test.cc:64-97
    assert shorten(ikey(b"foo", 100, 1), ikey(b"foo", 99, 1)) == ikey(b"foo", 100, 1)
    assert shorten(ikey(b"foo", 100, 1), ikey(b"foo", 101, 1)) == ikey(b"foo", 100, 1)
    assert shorten(ikey(b"foo", 100, 1), ikey(b"foo", 100, 1)) == ikey(b"foo", 100, 1)
    assert shorten(ikey(b"foo", 100, 1), ikey(b"foo", 100, 0)) == ikey(b"foo", 100, 1)
    assert shorten(ikey(b"foo", 100, 1), ikey(b"bar", 99, 1)) == ikey(b"foo", 100, 1)
    assert shorten(ikey(b"foo", 100, 1), ikey(b"hello", 200, 1)) == ikey(b"g", MAXSEQ, 1)
    assert shorten(ikey(b"foo", 100, 1), ikey(b"foobar", 200, 1)) == ikey(b"foo", 100, 1)
    assert shorten(ikey(b"foobar", 100, 1), ikey(b"foo", 200, 1)) == ikey(b"foobar", 100, 1)


def test_internal_key_separator_edge_cases():
    # util/comparator.cc:29-48: the increment must stay below the limit byte,
    # and the result must be physically shorter (db/dbformat.cc:76-78).
    assert shorten(ikey(b"abc", 5, 1), ikey(b"abe", 9, 1)) == ikey(b"abc", 5, 1)  # same length: kept
    assert shorten(ikey(b"abcd", 5, 1), ikey(b"abe", 9, 1)) == ikey(b"abd", MAXSEQ, 1)  # c+1 < e: shortened
    assert shorten(ikey(b"abcd", 5, 1), ikey(b"abd", 9, 1)) == ikey(b"abcd", 5, 1)  # c+1 == d: kept
    assert shorten(ikey(b"abcx", 5, 1), ikey(b"abez", 9, 1)) == ikey(b"abd", MAXSEQ, 1)
    assert shorten(ikey(b"ab\xffx", 5, 1), ikey(b"ac", 9, 1)) == ikey(b"ab\xffx", 5, 1)
    assert shorten(ikey(b"a\x01zz", 5, 1), ikey(b"a\x02", 9, 1)) == ikey(b"a\x01zz", 5, 1)


def test_internal_key_shortest_successor():
    # db/dbformat_test.cc:99-104
    assert successor(ikey(b"foo", 100, 1)) == ikey(b"g", MAXSEQ, 1)
    assert successor(ikey(b"\xff\xff", 100, 1)) == ikey(b"\xff\xff", 100, 1)
    assert successor(ikey(b"\xff\x01\x05", 7, 0)) == ikey(b"\xff\x02", MAXSEQ, 1)
    assert successor(ikey(b"\xfe", 7, 0)) == ikey(b"\xfe", 7, 0)  # not shorter: kept


def test_internal_key_compare_order():
    # db/dbformat.cc:51-67: user key ascending, then tag descending
    c = lambda a, b: O.or_ikey_compare(a, len(a), b, len(b))
    assert c(ikey(b"a", 5, 1), ikey(b"b", 1, 1)) < 0
    assert c(ikey(b"a", 5, 1), ikey(b"a", 4, 1)) < 0
    assert c(ikey(b"a", 5, 1), ikey(b"a", 5, 0)) < 0
    assert c(ikey(b"ab", 1, 1), ikey(b"a", 9, 1)) > 0
    assert c(ikey(b"\x80", 1, 1), ikey(b"\x7f", 1, 1)) > 0  # bytes compare unsigned


class Filter:
    def __init__(self, policy=1, bits=10):
        self.p = policy
        self.f = O.or_filter_new(policy, bits)

    def start(self, off):
        O.or_filter_start_block(self.f, off)

    def add(self, k):
        O.or_filter_add_key(self.f, k, len(k))

    def finish(self):
        n = C.c_size_t()
        p = O.or_filter_finish(self.f, C.byref(n))
        b = C.string_at(p, n.value)
        O.or_filter_free(self.f)
        return b


def may(policy, block, off, key):
    return O.or_filter_key_may_match(policy, block, len(block), off, key, len(key))


def test_filter_block_empty_builder():
    # table/filter_block_test.cc:44-51
    b = Filter().finish()
    assert b == b"\x00\x00\x00\x00\x0b"
    assert may(1, b, 0, b"foo") and may(1, b, 100000, b"foo")


def test_filter_block_single_chunk():
    # table/filter_block_test.cc:53-71
    f = Filter()
    f.start(100)
    for k in (b"foo", b"bar", b"box"):
        f.add(k)
    f.start(200)
    f.add(b"box")
    f.start(300)
    f.add(b"hello")
    b = f.finish()
    for k in (b"foo", b"bar", b"box", b"hello", b"foo"):
        assert may(1, b, 100, k)
    assert not may(1, b, 100, b"missing") and not may(1, b, 100, b"other")


def test_filter_block_multi_chunk():
    # table/filter_block_test.cc:73-118
    f = Filter()
    f.start(0)
    f.add(b"foo")
    f.start(2000)
    f.add(b"bar")
    f.start(3100)
    f.add(b"box")
    f.start(9000)
    f.add(b"box")
    f.add(b"hello")
    b = f.finish()
    assert may(1, b, 0, b"foo") and may(1, b, 2000, b"bar")
    assert not may(1, b, 0, b"box") and not may(1, b, 0, b"hello")
    assert may(1, b, 3100, b"box")
    assert not any(may(1, b, 3100, k) for k in (b"foo", b"bar", b"hello"))
    assert not any(may(1, b, 4100, k) for k in (b"foo", b"bar", b"box", b"hello"))
    assert may(1, b, 9000, b"box") and may(1, b, 9000, b"hello")
    assert not may(1, b, 9000, b"foo") and not may(1, b, 9000, b"bar")
    # layout closed form: 5 filters (windows 0..4), window 0 holds 2 hashes,
    # window 1 one, windows 2-3 empty, window 4 two; offsets array then 0x0b.
    data_len = 4 * 5
    assert len(b) == data_len + 5 * 4 + 4 + 1
    offs = struct.unpack_from("<5I", b, data_len)
    assert offs == (0, 8, 12, 12, 12)
    assert struct.unpack_from("<I", b, len(b) - 5)[0] == data_len


def bloom(keys, bits=10):
    kb = b"".join(keys)
    off = (C.c_uint64 * (len(keys) + 1))()
    o = 0
    for i, k in enumerate(keys):
        off[i] = o
        o += len(k)
    off[len(keys)] = o
    dst = C.create_string_buffer(len(keys) * 2 + 200)
    n = O.or_bloom_create(kb, off, len(keys), bits, dst, len(dst))
    return dst.raw[:n]


def test_bloom_closed_form_and_behaviour():
    # util/bloom.cc:28-54: bits = max(64, n*10) rounded to bytes, k = 6 byte
    assert len(bloom([])) == 8 + 1 and bloom([])[-1] == 6
    assert bloom([]) == bytes(8) + b"\x06"
    f = bloom([b"hello", b"world"])
    assert len(f) == 9 and f[-1] == 6
    assert O.or_bloom_may_match(b"hello", 5, f, len(f)) and O.or_bloom_may_match(b"world", 5, f, len(f))
    assert not O.or_bloom_may_match(b"x", 1, f, len(f)) and not O.or_bloom_may_match(b"foo", 3, f, len(f))
    # one key: exactly the k=6 double-hashed bits are set (bloom.cc:44-51)
    k = b"abc"
    f = bloom([k])
    h = O.or_hash(k, 3, 0xbc9f1d34)
    d = ((h >> 17) | (h << 15)) & 0xffffffff
    want = bytearray(8)
    for _ in range(6):
        bp = h % 64
        want[bp // 8] |= 1 << (bp % 8)
        h = (h + d) & 0xffffffff
    assert f[:8] == bytes(want)
    # util/bloom_test.cc:83-150: size bound and false-positive rate
    for n in (10, 100, 1000, 10000):
        keys = [struct.pack("<I", i) for i in range(n)]
        f = bloom(keys)
        assert len(f) <= n * 10 // 8 + 40
        assert all(O.or_bloom_may_match(x, 4, f, len(f)) for x in keys)
        fp = sum(O.or_bloom_may_match(struct.pack("<I", i + 1000000000), 4, f, len(f)) for i in range(10000))
        assert fp / 10000 <= 0.02

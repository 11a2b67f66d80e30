// dev_common.h — device-side building blocks shared by every kernel of the
// engine: byte coding (util/coding.cc), CRC-32C (util/crc32c.cc), the bloom
// hash (util/hash.cc), and the key-plane representation of internal keys
// (db/dbformat.h:165-175) with its comparator (db/dbformat.cc:51-67).
//
// Key planes.  Every entry's user key is stored as W 64-bit words
// P[w * stride + e], word w holding user-key bytes [8w, 8w+8) in big-endian
// order and zero-padded past the key's length; klen[e] holds the length and
// tag[e] the fixed64 tag (seq << 8 | type).  Comparing words as integers,
// then lengths, then tags descending gives InternalKeyComparator order for
// any keys of at most 8W bytes (a zero-padded prefix ties on the words and
// loses on the length, as memcmp-then-length does).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;

// Error bits in DevStatus::err.
#define ERR_CORRUPT 1u      // malformed table bytes: LUDA_CORRUPTION
#define ERR_UNSUPPORTED 2u  // valid but outside what the engine reproduces exactly
#define ERR_OVERFLOW 4u     // a capacity guess was too small: the host retries

// Where an error was first seen (DevStatus::site), for messages.
enum ErrSite : u32 {
  SITE_NONE = 0,
  SITE_FOOTER = 1,
  SITE_INDEX = 2,
  SITE_BLOCK_TRAILER = 3,
  SITE_ENTRY = 4,
  SITE_GROUP_COUNT = 5,
  SITE_SHORT_KEY = 6,
  SITE_KEY_WIDTH = 7,
  SITE_UNSORTED = 8,
  SITE_CAPACITY = 9,
  SITE_COMPRESSED = 10,
  SITE_BLOCK_ENTRIES = 11,
  SITE_FILE_BLOCKS = 12,
  SITE_SCAN = 13,
  SITE_CHECKSUM = 14,
};

// Device-resident counters and flags of one job.  The host reads this struct
// back twice per job: after the index/restart scan and at the end.
struct DevStatus {
  u32 err;
  u32 site;
  u32 max_ulen;       // longest user key seen while decoding
  u32 max_rulen;      // longest user key at a restart point (first guess of the key width)
  u32 max_blk;        // largest input data block, bytes
  u32 n_blocks_in;    // data blocks of all inputs
  u32 n_entries;      // entries of all inputs
  u32 n_kept;         // entries that survive the drop rule
  u32 n_blocks_out;
  u32 n_files_out;
  u32 n_windows;      // filter windows of all outputs
  u32 blk_step_max;   // most entries in one output block
  u32 file_step_max;  // most blocks in one output file
  u32 n_stops;        // ShouldStopBefore cuts that fired
  u32 n_forced;       // output blocks that start a table because of such a cut
  u32 need_files;     // tables, when over the host's capacity
  u32 need_windows;   // filter windows, when over the host's capacity
  u32 max_nr;         // most restart points in one input data block
  u32 need_blocks;    // output blocks, when over the host's capacity
  u32 max_ent;        // most entries in one input data block
  u64 out_bytes;      // sum of output table sizes
  u64 data_bytes;     // output data blocks with trailers
  u64 keys_needed;    // bytes of smallest/largest keys
  u32 bad_file;       // luda_verify_tables: lowest table with a failed check
  u32 cut_fallback;   // the tiled block cuts cannot take this stream: run the orbit kernels
};

// ---------------------------------------------------------------------------
// byte coding, util/coding.cc

__device__ __forceinline__ u32 ld32(const u8* p) {
  return (u32)p[0] | ((u32)p[1] << 8) | ((u32)p[2] << 16) | ((u32)p[3] << 24);
}
__device__ __forceinline__ u64 ld64(const u8* p) { return (u64)ld32(p) | ((u64)ld32(p + 4) << 32); }

// GetVarint32Ptr (util/coding.cc:121-140): nullptr on overrun or >5 bytes.
__device__ __forceinline__ const u8* dec_varint32(const u8* p, const u8* lim, u32* v) {
  u32 r = 0;
  for (u32 shift = 0; shift <= 28 && p < lim; shift += 7) {
    u32 b = *p++;
    if (b & 128) {
      r |= (b & 127) << shift;
    } else {
      r |= b << shift;
      *v = r;
      return p;
    }
  }
  return nullptr;
}
// GetVarint64Ptr (util/coding.cc:150-165)
__device__ __forceinline__ const u8* dec_varint64(const u8* p, const u8* lim, u64* v) {
  u64 r = 0;
  for (u32 shift = 0; shift <= 63 && p < lim; shift += 7) {
    u64 b = *p++;
    if (b & 128) {
      r |= (b & 127) << shift;
    } else {
      r |= b << shift;
      *v = r;
      return p;
    }
  }
  return nullptr;
}
__host__ __device__ __forceinline__ u32 varint_len(u64 v) {
  u32 n = 1;
  while (v >= 128) {
    v >>= 7;
    n++;
  }
  return n;
}
__device__ __forceinline__ u8* enc_varint(u8* p, u64 v) {
  while (v >= 128) {
    *p++ = (u8)(v | 128);
    v >>= 7;
  }
  *p++ = (u8)v;
  return p;
}
__device__ __forceinline__ void st32(u8* p, u32 v) {
  p[0] = (u8)v;
  p[1] = (u8)(v >> 8);
  p[2] = (u8)(v >> 16);
  p[3] = (u8)(v >> 24);
}

// ---------------------------------------------------------------------------
// CRC-32C (Castagnoli), util/crc32c.cc:276-377.  Tables are derived at
// compile time from the reflected polynomial 0x82F63B78.

#define CRC_ZERO_LEVELS 22  // shifts by 4, 8, 16, ..., 8 MiB zero bytes

// CRC-32C on a wavefront.  The message is left-padded with zeros to whole
// rows of 256 bytes; lane i folds the 4-byte words at row offsets 4i with
// slicing-by-4 lookups and a "append 256 zero bytes" step between rows; a
// 6-level tree then joins the lanes with "append 4*2^j zero bytes" steps.
// The ~0 initial register is folded into the message by complementing its
// first four bytes (a zero-initialised register fed X equals a ~0-initialised
// one fed X with its first 32 bits flipped), so zero padding is neutral.
struct CrcTables {
  u32 slice[8][256];                   // slice[k][b]: byte b followed by k zero bytes, from a zero register
  u32 zeros[CRC_ZERO_LEVELS][4][256];  // zeros[j]: "append 4*2^j zero bytes", one table per register byte
  constexpr CrcTables() : slice(), zeros() {
    for (u32 i = 0; i < 256; i++) {
      u32 c = i;
      for (int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ 0x82F63B78u : (c >> 1);
      slice[0][i] = c;
    }
    for (int k = 1; k < 8; k++)
      for (u32 i = 0; i < 256; i++) {
        u32 s = slice[k - 1][i];
        slice[k][i] = slice[0][s & 0xff] ^ (s >> 8);
      }
    // The register update is linear over GF(2), so "feed n zero bytes" is a
    // 32x32 bit matrix, kept as its 32 columns.  Level 0 (4 bytes) by direct
    // feeding, level j+1 as level j applied twice.
    u32 col[32] = {};
    for (int bit = 0; bit < 32; bit++) {
      u32 s = 1u << bit;
      for (int k = 0; k < 4; k++) s = slice[0][s & 0xff] ^ (s >> 8);
      col[bit] = s;
    }
    for (int j = 0; j < CRC_ZERO_LEVELS; j++) {
      for (int b = 0; b < 4; b++)
        for (u32 v = 0; v < 256; v++) {
          u32 r = 0;
          for (int bit = 0; bit < 8; bit++)
            if (v & (1u << bit)) r ^= col[8 * b + bit];
          zeros[j][b][v] = r;
        }
      u32 next[32] = {};
      for (int bit = 0; bit < 32; bit++) {
        u32 x = col[bit], y = 0;
        for (int t = 0; t < 32; t++)
          if (x & (1u << t)) y ^= col[t];
        next[bit] = y;
      }
      for (int bit = 0; bit < 32; bit++) col[bit] = next[bit];
    }
  }
};

// One copy per translation unit (device code is not linked across units).
static __device__ const CrcTables g_crc = CrcTables();

// Apply a zero-append map (one of g_crc.zeros) to a raw register value.
__device__ __forceinline__ u32 crc_shift(const u32 (*z)[256], u32 s) {
  return z[0][s & 0xff] ^ z[1][(s >> 8) & 0xff] ^ z[2][(s >> 16) & 0xff] ^ z[3][s >> 24];
}

// crc32c::Value of a message of L >= 4 bytes, byte(i) returning byte i.
// Called by all 64 lanes of a wavefront; every lane gets the result.
// `slice` and `z256` may point at LDS copies of g_crc.slice / g_crc.zeros[6].
// Raw CRC-32C register (zero start, no final xor) of rows [r0, r1) of the
// message of L >= 4 bytes left-padded to whole 256-byte rows, with the ~0
// initial value folded into the message's first four bytes.  Called by all
// 64 lanes of a wavefront; every lane gets the result.
template <typename ByteAt>
__device__ __forceinline__ u32 wave_crc32c_rows(u64 L, u64 r0, u64 r1, const u32 (*slice)[256],
                                                const u32 (*z256)[256], ByteAt at) {
  const u32 lane = threadIdx.x & 63;
  const u64 P = (256 - (L & 255)) & 255;
  u32 s = 0;
  // Rows go in batches whose bytes are all fetched before any is folded, so
  // a message in global memory costs one load round trip per batch.
  constexpr u32 kRows = 8;
  for (u64 rb = r0; rb < r1; rb += kRows) {
    u32 b[kRows][4];
#pragma unroll
    for (u32 rr = 0; rr < kRows; rr++) {
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const long long x = (long long)(256 * (rb + rr) + 4 * lane + t) - (long long)P;
        const bool in = rb + rr < r1 && x >= 0 && (u64)x < L;
        // The index is clamped before it reaches an address: a padding byte
        // reads byte 0 and is then zeroed.  No address below the message is
        // ever formed, not even as the base of a folded immediate offset.
        const u64 xi = in ? (u64)x : 0;
        u32 v = (u32)at(xi);
        v = in ? v : 0u;
        if (x >= 0 && x < 4) v ^= 0xffu;
        b[rr][t] = v;
      }
    }
#pragma unroll
    for (u32 rr = 0; rr < kRows; rr++) {
      if (rb + rr < r1) {
        const u32 w = slice[3][b[rr][0]] ^ slice[2][b[rr][1]] ^ slice[1][b[rr][2]] ^ slice[0][b[rr][3]];
        s = crc_shift(z256, s) ^ w;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 6; j++) {
    u32 other = __shfl_down(s, 1u << j, 64);
    if ((lane & ((2u << j) - 1)) == 0) s = crc_shift(g_crc.zeros[j], s) ^ other;
  }
  return __shfl(s, 0, 64);
}

// wave_crc32c_rows for a message held in LDS at byte `mis` of the 4-byte
// words s32 (readable for 4 bytes past the message): every row but the ones
// touching the padding or the first four bytes reads its four bytes per
// lane with two aligned LDS words and a byte-align.
__device__ __forceinline__ u32 wave_crc32c_rows_lds(u64 L, u64 r0, u64 r1, const u32* s32, u32 mis,
                                                    const u32 (*slice)[256], const u32 (*z256)[256]) {
  const u32 lane = threadIdx.x & 63;
  const u64 P = (256 - (L & 255)) & 255;
  const u8* s8 = reinterpret_cast<const u8*>(s32);
  u32 s = 0;
  constexpr u32 kRows = 8;
  for (u64 rb = r0; rb < r1; rb += kRows) {
    u32 wv[kRows];
#pragma unroll
    for (u32 rr = 0; rr < kRows; rr++) {
      const long long x = (long long)(256 * (rb + rr) + 4 * lane) - (long long)P;
      u32 v = 0;
      if (rb + rr < r1) {
        if (x >= 4 && (u64)x + 4 <= L) {
          const u32 o = mis + (u32)x;
          v = __builtin_amdgcn_alignbyte(s32[(o >> 2) + 1], s32[o >> 2], o & 3);
        } else {
#pragma unroll
          for (int t = 0; t < 4; t++) {
            const long long xt = x + t;
            const bool in = xt >= 0 && (u64)xt < L;
            u32 bt = in ? (u32)s8[mis + (u32)(in ? xt : 0)] : 0u;
            if (xt >= 0 && xt < 4) bt ^= 0xffu;
            v |= bt << (8 * t);
          }
        }
      }
      wv[rr] = v;
    }
#pragma unroll
    for (u32 rr = 0; rr < kRows; rr++) {
      if (rb + rr < r1) {
        const u32 v = wv[rr];
        const u32 w = slice[3][v & 0xff] ^ slice[2][(v >> 8) & 0xff] ^ slice[1][(v >> 16) & 0xff] ^ slice[0][v >> 24];
        s = crc_shift(z256, s) ^ w;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 6; j++) {
    u32 other = __shfl_down(s, 1u << j, 64);
    if ((lane & ((2u << j) - 1)) == 0) s = crc_shift(g_crc.zeros[j], s) ^ other;
  }
  return __shfl(s, 0, 64);
}

// crc_append_zeros with the first nz levels' maps read from zt (copies of
// g_crc.zeros[0 .. nz)), the longer ones from g_crc.
__device__ __forceinline__ u32 crc_append_zeros_tab(u32 s, u64 n, const u32 (*zt)[4][256], u32 nz) {
  u64 w = n >> 2;
  for (u32 j = 0; j < CRC_ZERO_LEVELS && w; j++, w >>= 1)
    if (w & 1) s = j < nz ? crc_shift(zt[j], s) : crc_shift(g_crc.zeros[j], s);
  return s;
}

// Raw register s followed by n zero bytes (n a multiple of 4, < 16 MiB).
__device__ __forceinline__ u32 crc_append_zeros(u32 s, u64 n) {
  u64 w = n >> 2;
  for (int j = 0; j < CRC_ZERO_LEVELS && w; j++, w >>= 1)
    if (w & 1) s = crc_shift(g_crc.zeros[j], s);
  return s;
}

// crc32c::Value of a message of L >= 4 bytes, byte(i) returning byte i.
// Called by all 64 lanes of a wavefront; every lane gets the result.
// `slice` and `z256` may point at LDS copies of g_crc.slice / g_crc.zeros[6].
template <typename ByteAt>
__device__ __forceinline__ u32 wave_crc32c(u64 L, const u32 (*slice)[256], const u32 (*z256)[256], ByteAt at) {
  const u64 rows = (L + 255) >> 8;
  return wave_crc32c_rows(L, 0, rows, slice, z256, at) ^ 0xffffffffu;
}

// util/crc32c.h:29-32
__host__ __device__ __forceinline__ u32 crc_mask(u32 crc) { return ((crc >> 15) | (crc << 17)) + 0xa282ead8u; }

// ---------------------------------------------------------------------------
// util/hash.cc:22-53 over bytes produced by a getter.

template <typename ByteAt>
__device__ __forceinline__ u32 leveldb_hash(u32 n, u32 seed, ByteAt at) {
  const u32 m = 0xc6a4a793u;
  u32 h = seed ^ (n * m);
  u32 i = 0;
  for (; i + 4 <= n; i += 4) {
    u32 w = (u32)at(i) | ((u32)at(i + 1) << 8) | ((u32)at(i + 2) << 16) | ((u32)at(i + 3) << 24);
    h += w;
    h *= m;
    h ^= (h >> 16);
  }
  switch (n - i) {
    case 3:
      h += (u32)at(i + 2) << 16;
      [[fallthrough]];
    case 2:
      h += (u32)at(i + 1) << 8;
      [[fallthrough]];
    case 1:
      h += at(i);
      h *= m;
      h ^= (h >> 24);
      break;
  }
  return h;
}

// ---------------------------------------------------------------------------
// key planes

struct Planes {
  u64* P;       // W planes of `stride` words
  u32* klen;    // user key length
  u64* tag;     // seq << 8 | type
  u64 stride;
  u32 W;
};

__device__ __forceinline__ u8 ukey_byte(const Planes& k, u64 e, u32 i) {
  if ((i >> 3) >= k.W) return 0;  // only reachable on a rejected job; keeps reads in bounds
  u64 w = k.P[(u64)(i >> 3) * k.stride + e];
  return (u8)(w >> (56 - 8 * (i & 7)));
}
// byte i of the internal key user_key || fixed64(tag)
__device__ __forceinline__ u8 ikey_byte(const Planes& k, u64 e, u32 ulen, u64 tag, u32 i) {
  return i < ulen ? ukey_byte(k, e, i) : (u8)(tag >> (8 * (i - ulen)));
}

// InternalKeyComparator::Compare for two entries of one plane set.
__device__ __forceinline__ int ikey_cmp(const Planes& k, u64 a, u64 b) {
  for (u32 w = 0; w < k.W; w++) {
    u64 x = k.P[w * k.stride + a], y = k.P[w * k.stride + b];
    if (x != y) return x < y ? -1 : 1;
  }
  u32 la = k.klen[a], lb = k.klen[b];
  if (la != lb) return la < lb ? -1 : 1;
  u64 ta = k.tag[a], tb = k.tag[b];
  if (ta != tb) return ta > tb ? -1 : 1;
  return 0;
}
__device__ __forceinline__ bool ukey_eq(const Planes& k, u64 a, u64 b) {
  if (k.klen[a] != k.klen[b]) return false;
  for (u32 w = 0; w < k.W; w++)
    if (k.P[w * k.stride + a] != k.P[w * k.stride + b]) return false;
  return true;
}

// A key held outside the planes (grandparent largest key, deeper-level
// range ends): its first 8W bytes as words, and its full length.  Keys longer
// than 8W bytes compare correctly against plane keys of at most 8W bytes.
struct FixedKey {
  u32 len;   // user-key length
  u32 pad;
  u64 tag;   // only for internal keys
};

// Bytewise compare of plane key e against fixed key f (words fw[0..W)).
__device__ __forceinline__ int ukey_cmp_fixed(const Planes& k, u64 e, const u64* fw, u32 flen) {
  for (u32 w = 0; w < k.W; w++) {
    u64 x = k.P[w * k.stride + e], y = fw[w];
    if (x != y) return x < y ? -1 : 1;
  }
  u32 la = k.klen[e];
  if (la != flen) return la < flen ? -1 : 1;
  return 0;
}

// Global-memory views of device addresses that kernels get as integers or
// load from memory (value references, table and block addresses).  Reading
// through an address-space-1 pointer tells the compiler these bytes are not
// in LDS, so the loads are neither generic (flat) nor ordered behind the
// kernel's LDS traffic.  The unaligned types allow any byte offset.
#define GMEM __attribute__((address_space(1)))
typedef u32 __attribute__((aligned(1))) u32_una;
typedef u64 __attribute__((aligned(1))) u64_una;
__device__ __forceinline__ const GMEM u8* gbytes(u64 addr) { return (const GMEM u8*)addr; }
__device__ __forceinline__ u32 gload_u32(const GMEM u8* p) { return *(const GMEM u32_una*)p; }
__device__ __forceinline__ u64 gload_u64(const GMEM u8* p) { return *(const GMEM u64_una*)p; }

// wave_crc32c_rows for a message at p in global memory, readable for its L
// bytes only: rows away from the ends read their four bytes per lane with
// one (unaligned) dword load, the others byte by byte at clamped indices.
// zt[j] (j < 6) are copies of g_crc.zeros[j] for the join of the lanes.
__device__ __forceinline__ u32 wave_crc32c_rows_global(u64 L, u64 r0, u64 r1, const GMEM u8* p,
                                                       const u32 (*slice)[256], const u32 (*z256)[256],
                                                       const u32 (*zt)[4][256]) {
  const u32 lane = threadIdx.x & 63;
  const u64 P = (256 - (L & 255)) & 255;
  u32 s = 0;
  constexpr u32 kRows = 8;
  for (u64 rb = r0; rb < r1; rb += kRows) {
    u32 wv[kRows];
#pragma unroll
    for (u32 rr = 0; rr < kRows; rr++) {
      const long long x = (long long)(256 * (rb + rr) + 4 * lane) - (long long)P;
      u32 v = 0;
      if (rb + rr < r1) {
        if (x >= 4 && (u64)x + 4 <= L) {
          v = gload_u32(p + x);
        } else {
#pragma unroll
          for (int t = 0; t < 4; t++) {
            const long long xt = x + t;
            const bool in = xt >= 0 && (u64)xt < L;
            u32 bt = in ? (u32)p[in ? xt : 0] : 0u;
            if (xt >= 0 && xt < 4) bt ^= 0xffu;
            v |= bt << (8 * t);
          }
        }
      }
      wv[rr] = v;
    }
#pragma unroll
    for (u32 rr = 0; rr < kRows; rr++) {
      if (rb + rr < r1) {
        const u32 v = wv[rr];
        const u32 w = slice[3][v & 0xff] ^ slice[2][(v >> 8) & 0xff] ^ slice[1][(v >> 16) & 0xff] ^ slice[0][v >> 24];
        s = crc_shift(z256, s) ^ w;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 6; j++) {
    u32 other = __shfl_down(s, 1u << j, 64);
    if ((lane & ((2u << j) - 1)) == 0) s = crc_shift(zt[j], s) ^ other;
  }
  return __shfl(s, 0, 64);
}

// ---------------------------------------------------------------------------
// Wavefront scans on the VALU.  DPP moves a value between lanes without the
// LDS crossbar that __shfl (ds_bpermute) uses: row_shr:1,2,4,8 scan within
// rows of 16 lanes, row_bcast:15 and row_bcast:31 carry the row totals.  A
// lane with no source (or in a row the step leaves alone) reads 0, which is
// the identity of every operator scanned here.

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ u32 dpp_u32(u32 v) {
  return (u32)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ u64 dpp_u64(u64 v) {
  return (u64)dpp_u32<CTRL, ROW_MASK>((u32)v) | ((u64)dpp_u32<CTRL, ROW_MASK>((u32)(v >> 32)) << 32);
}

// Inclusive scan over the 64 lanes with an associative operator
// op(earlier, later); `Mov` moves one lane's state (DPP on each word).
template <typename S, typename Op>
__device__ __forceinline__ S wave_incl_scan_op(S x, Op op) {
  x = op(x.template dpp<0x111, 0xf>(), x);  // row_shr:1
  x = op(x.template dpp<0x112, 0xf>(), x);  // row_shr:2
  x = op(x.template dpp<0x114, 0xf>(), x);  // row_shr:4
  x = op(x.template dpp<0x118, 0xf>(), x);  // row_shr:8
  x = op(x.template dpp<0x142, 0xa>(), x);  // row_bcast:15 into rows 1 and 3
  x = op(x.template dpp<0x143, 0xc>(), x);  // row_bcast:31 into rows 2 and 3
  return x;
}

// wave_incl_scan_op correct in lanes [0, 2^STEPS) only (STEPS = 3, 4, 5):
// the steps that carry into the lanes past them are left out.
template <u32 STEPS, typename S, typename Op>
__device__ __forceinline__ S wave_incl_scan_op_low(S x, Op op) {
  x = op(x.template dpp<0x111, 0xf>(), x);  // row_shr:1
  x = op(x.template dpp<0x112, 0xf>(), x);  // row_shr:2
  x = op(x.template dpp<0x114, 0xf>(), x);  // row_shr:4
  if constexpr (STEPS >= 4) x = op(x.template dpp<0x118, 0xf>(), x);  // row_shr:8
  if constexpr (STEPS >= 5) x = op(x.template dpp<0x142, 0xa>(), x);  // row_bcast:15 into rows 1 and 3
  return x;
}

__device__ __forceinline__ u32 wave_incl_sum(u32 v) {
  v += dpp_u32<0x111, 0xf>(v);
  v += dpp_u32<0x112, 0xf>(v);
  v += dpp_u32<0x114, 0xf>(v);
  v += dpp_u32<0x118, 0xf>(v);
  v += dpp_u32<0x142, 0xa>(v);
  v += dpp_u32<0x143, 0xc>(v);
  return v;
}
__device__ __forceinline__ u32 wave_incl_max(u32 v) {
  v = max(v, dpp_u32<0x111, 0xf>(v));
  v = max(v, dpp_u32<0x112, 0xf>(v));
  v = max(v, dpp_u32<0x114, 0xf>(v));
  v = max(v, dpp_u32<0x118, 0xf>(v));
  v = max(v, dpp_u32<0x142, 0xa>(v));
  v = max(v, dpp_u32<0x143, 0xc>(v));
  return v;
}
__device__ __forceinline__ u64 wave_incl_sum(u64 v) {
  v += dpp_u64<0x111, 0xf>(v);
  v += dpp_u64<0x112, 0xf>(v);
  v += dpp_u64<0x114, 0xf>(v);
  v += dpp_u64<0x118, 0xf>(v);
  v += dpp_u64<0x142, 0xa>(v);
  v += dpp_u64<0x143, 0xc>(v);
  return v;
}
// Lane 63's value in every lane (a scalar read, no LDS).
__device__ __forceinline__ u32 wave_last(u32 v) { return (u32)__builtin_amdgcn_readlane((int)v, 63); }
__device__ __forceinline__ u64 wave_last(u64 v) {
  return (u64)(u32)__builtin_amdgcn_readlane((int)(u32)v, 63) |
         ((u64)(u32)__builtin_amdgcn_readlane((int)(u32)(v >> 32), 63) << 32);
}

// Grid-stride helpers.
#define GRID_STRIDE(i, n) for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < (u64)(n); i += (u64)gridDim.x * blockDim.x)

__device__ __forceinline__ void set_err(DevStatus* st, u32 bits, u32 site) {
  atomicOr(&st->err, bits);
  atomicCAS(&st->site, 0u, site);
}

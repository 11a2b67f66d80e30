// k_encode.hip — the table writer: TableBuilder (table/table_builder.cc),
// BlockBuilder in the fork's format (table/block_builder.cc), the filter
// block (table/filter_block.cc) with the bloom policy (util/bloom.cc) over
// user keys (db/dbformat.cc:105-115), block trailers with masked CRC-32C
// (table_builder.cc:178-195), the metaindex and index blocks with shortened
// separators (db/dbformat.cc:69-101) and the footer (table/format.cc:31-40),
// plus DoCompactionWork's cut into tables (db/db_impl.cc:1158-1181).
//
// Input: the kept stream (entries in output order, `forced` where a table
// must start).  Every size the sequential writer would compute one entry at a
// time is computed per entry, block or table in parallel; the two greedy cut
// sequences (blocks, tables) come from the orbit kernels.
#include <algorithm>

#include "engine.h"

namespace {

constexpr u32 kThreads = 256;
constexpr u32 kWaves = kThreads / 64;
constexpr u32 kStage = 6144;   // bytes of one data block assembled in LDS per wave
constexpr u32 kFBits = 4096;   // bytes of one bloom filter assembled in LDS per wave
constexpr u64 kMaxSeqTag = ((((u64)1 << 56) - 1) << 8) | 1;  // kMaxSequenceNumber, kValueTypeForSeek

__device__ __forceinline__ void wave_sync() {
  __threadfence_block();
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ u32 wave_incl_scan(u32 v) { return wave_incl_sum(v); }

__device__ __forceinline__ u32 wave_max(u32 v) {
#pragma unroll
  for (u32 d = 32; d > 0; d >>= 1) v = max(v, (u32)__shfl_xor(v, d, 64));
  return v;
}

// First byte where the user keys of a and b differ, capped at min length.
__device__ __forceinline__ u32 ukey_diff(const Planes& K, u64 a, u64 b, u32 m) {
  for (u32 w = 0; w < K.W; w++) {
    const u64 x = K.P[w * K.stride + a] ^ K.P[w * K.stride + b];
    if (x) {
      const u32 pos = 8 * w + ((u32)__clzll((long long)x) >> 3);
      return pos < m ? pos : m;
    }
  }
  return m;
}

// Shared prefix of the internal keys user_key || fixed64(tag) of a and b
// (BlockBuilder::Add, table/block_builder.cc:80-87).
__device__ u32 ikey_lcp(const Planes& K, u64 a, u64 b) {
  const u32 la = K.klen[a], lb = K.klen[b];
  const u32 m = min(la, lb);
  const u32 d = ukey_diff(K, a, b, m);
  if (d < m) return d;
  const u64 ta = K.tag[a], tb = K.tag[b];
  const u32 lim = min(la, lb) + 8;
  u32 i = m;
  while (i < lim && ikey_byte(K, a, la, ta, i) == ikey_byte(K, b, lb, tb, i)) i++;
  return i;
}

// Index key for a block boundary: InternalKeyComparator::FindShortestSeparator
// (db/dbformat.cc:69-86) over BytewiseComparatorImpl's (util/comparator.cc:29-48).
// Returns the key length; writes the key when out is non-null.
__device__ u32 index_sep(const Planes& K, u64 a, u64 b, u8* out) {
  const u32 la = K.klen[a], lb = K.klen[b];
  const u32 m = min(la, lb);
  const u32 d = ukey_diff(K, a, b, m);
  if (d < m) {
    const u32 da = ukey_byte(K, a, d), db = ukey_byte(K, b, d);
    if (da < 0xff && da + 1 < db && d + 1 < la) {
      if (out) {
        for (u32 i = 0; i < d; i++) out[i] = ukey_byte(K, a, i);
        out[d] = (u8)(da + 1);
        for (u32 i = 0; i < 8; i++) out[d + 1 + i] = (u8)(kMaxSeqTag >> (8 * i));
      }
      return d + 1 + 8;
    }
  }
  if (out) {
    const u64 ta = K.tag[a];
    for (u32 i = 0; i < la + 8; i++) out[i] = ikey_byte(K, a, la, ta, i);
  }
  return la + 8;
}

// Index key after a table's last block: FindShortSuccessor
// (db/dbformat.cc:88-101, util/comparator.cc:50-63).
__device__ u32 index_succ(const Planes& K, u64 a, u8* out) {
  const u32 la = K.klen[a];
  for (u32 i = 0; i < la; i++) {
    const u32 c = ukey_byte(K, a, i);
    if (c != 0xff) {
      if (i + 1 < la) {
        if (out) {
          for (u32 t = 0; t < i; t++) out[t] = ukey_byte(K, a, t);
          out[i] = (u8)(c + 1);
          for (u32 t = 0; t < 8; t++) out[i + 1 + t] = (u8)(kMaxSeqTag >> (8 * t));
        }
        return i + 1 + 8;
      }
      break;
    }
  }
  if (out) {
    const u64 ta = K.tag[a];
    for (u32 i = 0; i < la + 8; i++) out[i] = ikey_byte(K, a, la, ta, i);
  }
  return la + 8;
}

// util/hash.cc over the user key of entry q, four bytes per step.
__device__ u32 hash_ukey(const Planes& K, u64 q, u32 seed) {
  const u32 n = min(K.klen[q], 8 * K.W);
  const u32 m = 0xc6a4a793u;
  u32 h = seed ^ (n * m);
  u32 i = 0;
  for (; i + 4 <= n; i += 4) {
    const u64 word = K.P[(u64)(i >> 3) * K.stride + q];
    const u32 half = (i & 4) ? (u32)word : (u32)(word >> 32);
    h += __builtin_bswap32(half);
    h *= m;
    h ^= (h >> 16);
  }
  switch (n - i) {
    case 3:
      h += (u32)ukey_byte(K, q, i + 2) << 16;
      [[fallthrough]];
    case 2:
      h += (u32)ukey_byte(K, q, i + 1) << 8;
      [[fallthrough]];
    case 1:
      h += ukey_byte(K, q, i);
      h *= m;
      h ^= (h >> 24);
      break;
  }
  return h;
}

// ---------------------------------------------------------------------------
// sizes and cut points

__global__ void k_entry_sizes(Kept k, const u32* m_dev, EncBufs b, DevStatus* st) {
  const u64 m = *m_dev;
  GRID_STRIDE(q, m) {
    const Planes& K = k.K;
    u32 sh = 0;
    if (q > 0) {
      if (ikey_cmp(K, q - 1, q) >= 0) set_err(st, ERR_UNSUPPORTED, SITE_UNSORTED);
      sh = ikey_lcp(K, q - 1, q);
    }
    const u32 il = K.klen[q] + 8;
    const u32 vl = k.vlen[q];
    b.shared[q] = sh;
    b.szr[q] = 1 + varint_len(il) + varint_len(vl) + il + vl;
    b.sznr[q] = varint_len(sh) + varint_len(il - sh) + varint_len(vl) + (il - sh) + vl;
  }
}

// The block that TableBuilder would build starting at entry q: entries are
// added until CurrentSizeEstimate() >= block_size (table_builder.cc:119-123),
// a table cut is due, or the stream ends.  A workgroup takes 256 starts and
// stages the sizes of the entries they can reach (kBnextSpan past the last
// start) in LDS; a walk that runs past the staged span reads on from global
// memory.
constexpr u32 kBnextSpan = 768;
constexpr u32 kBnextTile = 256 + kBnextSpan + 4;
__global__ __launch_bounds__(256) void k_bnext(Kept k, const u32* m_dev, EncParams ep, EncBufs b, DevStatus* st) {
  // per staged entry: szr | sznr << 31 | forced << 63 (sizes < 2^31, checked)
  __shared__ u64 s_pk[kBnextTile];
  const u64 m = *m_dev;
  const u32 R = ep.interval;
  u32 smax = 0;
  for (u64 q0 = (u64)blockIdx.x * 256; q0 < m; q0 += (u64)gridDim.x * 256) {
    const u64 tend = min(q0 + kBnextTile, m);  // staged entries [q0, tend)
    {
      constexpr u32 kPer = (kBnextTile + 255) / 256;
      u32 r[kPer], nr[kPer], f[kPer];
#pragma unroll
      for (u32 t = 0; t < kPer; t++) {
        const u64 j = q0 + threadIdx.x + 256 * t;
        const u64 jc = j < tend ? j : q0;
        r[t] = b.szr[jc];
        nr[t] = b.sznr[jc];
        f[t] = k.forced[jc];
      }
#pragma unroll
      for (u32 t = 0; t < kPer; t++) {
        const u32 x = threadIdx.x + 256 * t;
        if (x < kBnextTile) {
          if (r[t] >= (1u << 31) || nr[t] >= (1u << 31)) set_err(st, ERR_UNSUPPORTED, SITE_BLOCK_ENTRIES);
          s_pk[x] = q0 + x < tend ? ((u64)r[t] | ((u64)nr[t] << 31) | ((u64)(f[t] != 0) << 63)) : 0;
        }
      }
    }
    __syncthreads();
    const u64 q = q0 + threadIdx.x;
    if (q < m) {
      // acc: entry bytes so far; ng: restart points so far; rpos: position
      // of the next entry in its restart group (no division in the walk)
      u64 acc = 0, j = q;
      u32 ng = 0, rpos = 0;
      bool done = false;
      // within the staged span, four entries per LDS round trip ...
      while (!done && j + 5 <= tend) {
        const u32 x = (u32)(j - q0);
        u64 pk[5];
#pragma unroll
        for (u32 t = 0; t < 5; t++) pk[t] = s_pk[x + t];
#pragma unroll
        for (u32 t = 0; t < 4; t++) {
          if (!done) {
            const bool rs = rpos == 0;
            acc += rs ? (pk[t] & 0x7fffffffu) : ((pk[t] >> 31) & 0x7fffffffu);
            ng += rs;
            rpos = rpos + 1 == R ? 0 : rpos + 1;
            j++;
            done = acc + 4ull * ng + 4 >= ep.block_size || j >= m || (pk[t + 1] >> 63) != 0;
          }
        }
      }
      // ... then from global memory: the two never share a pointer
      while (!done) {
        const bool rs = rpos == 0;
        acc += rs ? b.szr[j] : b.sznr[j];
        ng += rs;
        rpos = rpos + 1 == R ? 0 : rpos + 1;
        j++;
        done = acc + 4ull * ng + 4 >= ep.block_size || j >= m || k.forced[j] != 0;
      }
      const u64 cnt = j - q;
      b.bnext[q] = (u32)j;
      b.bbytes[q] = (u32)(acc + 4ull * ng + 4);
      smax = max(smax, (u32)cnt);
    }
    __syncthreads();
  }
  smax = wave_max(smax);
  if ((threadIdx.x & 63) == 0 && smax) atomicMax(&st->blk_step_max, smax);
}

__global__ void k_block_info(Kept k, const u32* m_dev, EncBufs b, DevStatus* st) {
  const u64 m = *m_dev;
  const u32 nbo = st->n_blocks_out;  // 0 when the count overflowed its capacity (reported)
  GRID_STRIDE(q, m) {
    if (!b.bs[q]) continue;
    const u32 id = b.bid[q];
    if (id >= nbo) continue;
    b.ob_first[id] = (u32)q;
    b.ob_n[id] = b.bnext[q] - (u32)q;
    b.ob_bytes[id] = b.bbytes[q];
    b.ob_forced[id] = (q > 0 && k.forced[q]) ? 1u : 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    b.ob_forced[nbo] = 0;
    b.ob_bytes[nbo] = 0;
  }
}

// ob_bytes + 5 in place of a separate array for the offset scan.
__global__ void k_block_tot(const u32* nbo_dev, EncBufs b, u32* tot) {
  const u64 nbo = *nbo_dev;
  GRID_STRIDE(i, nbo + 1) tot[i] = i < nbo ? b.ob_bytes[i] + 5 : 0;
}

__global__ void k_block_forced(const u32* nbo_dev, EncBufs b) {
  const u64 nbo = *nbo_dev;
  GRID_STRIDE(i, nbo) {
    if (b.ob_forced[i]) b.flist[b.ff[i]] = (u32)i;
  }
}

// The table DoCompactionWork would build starting at block i: blocks are
// added until FileSize() >= max_file_size (db/db_impl.cc:1173-1180), or up to
// the next block that ShouldStopBefore starts (:1105-1111).
__global__ void k_fnext(const u32* nbo_dev, const u32* nforced_dev, EncParams ep, EncBufs b, DevStatus* st) {
  const u64 nbo = *nbo_dev;
  const u32 nforced = *nforced_dev;
  u32 smax = 0;
  GRID_STRIDE(i, nbo) {
    const u64 base = b.O[i];
    const u64 target = ep.max_file_size > ~0ull - base ? ~0ull : base + ep.max_file_size;
    u64 lo = i + 1, hi = nbo;  // smallest k in [i+1, nbo] with O[k] >= target, else nbo
    while (lo < hi) {
      const u64 mid = (lo + hi) >> 1;
      if (b.O[mid] >= target) hi = mid;
      else lo = mid + 1;
    }
    u64 nx = lo;
    const u32 before = b.ff[i] + b.ob_forced[i];  // forced blocks at or before i
    if (before < nforced) nx = min(nx, (u64)b.flist[before]);
    b.fnext[i] = (u32)nx;
    smax = max(smax, (u32)(nx - i));
  }
  smax = wave_max(smax);
  if ((threadIdx.x & 63) == 0 && smax) atomicMax(&st->file_step_max, smax);
}

__global__ void k_file_first(const u32* nbo_dev, EncBufs b, const u32* nf_dev) {
  const u64 nbo = *nbo_dev;
  const u32 nf = *nf_dev;  // 0 when the table count overflowed its capacity
  GRID_STRIDE(i, nbo) {
    if (b.fs[i] && b.fid[i] < nf) b.f_first[b.fid[i]] = (u32)i;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) b.f_first[*nf_dev] = (u32)nbo;
}

__device__ __forceinline__ u32 file_of_block(const EncBufs& b, u32 nf, u64 i) {
  u32 lo = 0, hi = nf;
  while (hi - lo > 1) {
    const u32 mid = (lo + hi) >> 1;
    if (b.f_first[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void k_block_layout(Kept k, const u32* nbo_dev, const u32* nf_dev, EncParams ep, EncBufs b) {
  const u64 nbo = *nbo_dev;
  const u32 nf = *nf_dev;
  if (nf == 0) return;  // no tables, or their count overflowed (reported)
  GRID_STRIDE(i, nbo) {
    const u32 f = file_of_block(b, nf, i);
    const u64 off = b.O[i] - b.O[b.f_first[f]];
    b.ob_file[i] = f;
    b.ob_off[i] = off;
    const u64 last = (u64)b.ob_first[i] + b.ob_n[i] - 1;
    const u32 kl = (i + 1 < b.f_first[f + 1]) ? index_sep(k.K, last, b.ob_first[i + 1], nullptr)
                                              : index_succ(k.K, last, nullptr);
    const u32 hl = varint_len(off) + varint_len(b.ob_bytes[i]);
    b.ob_isz[i] = 1 + varint_len(kl) + varint_len(hl) + kl + hl;
  }
}

__global__ void k_isz_sentinel(const u32* nbo_dev, EncBufs b) { b.ob_isz[*nbo_dev] = 0; }

// Filter windows of a table: FilterBlockBuilder emits one filter per 2 KiB of
// block start offsets up to the data end, plus one for keys still pending at
// Finish (table/filter_block.cc:21-49).
__global__ void k_file_layout(const u32* nf_dev, EncParams ep, EncBufs b) {
  const u64 nf = *nf_dev;
  GRID_STRIDE(f, nf) {
    const u32 b0 = b.f_first[f], b1 = b.f_first[f + 1];
    const u64 dend = b.O[b1] - b.O[b0];
    const u64 last = b.O[b1 - 1] - b.O[b0];
    b.f_dend[f] = dend;
    u64 nw = 0;
    if (ep.bloom_bits) nw = max(dend >> 11, (last >> 11) + 1);
    b.f_nwin[f] = (u32)nw;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) b.f_nwin[nf] = 0;
}

__global__ void k_window_init(const u32* nw_dev, EncBufs b) {
  const u64 nw = *nw_dev;
  GRID_STRIDE(w, nw + 1) {
    b.w_keys[w] = 0;
    b.w_first[w] = 0xffffffffu;
  }
}

__global__ void k_window_counts(const u32* nbo_dev, const u32* nf_dev, const u32* nw_dev, EncBufs b) {
  const u64 nbo = *nbo_dev;
  const u32 nw = *nw_dev;  // 0 when the window count overflowed its capacity
  if (*nf_dev == 0 || nw == 0) return;
  GRID_STRIDE(i, nbo) {
    const u32 w = b.f_win[b.ob_file[i]] + (u32)(b.ob_off[i] >> 11);
    if (w >= nw) continue;
    atomicAdd(&b.w_keys[w], b.ob_n[i]);
    atomicMin(&b.w_first[w], b.ob_first[i]);
  }
}

// BloomFilterPolicy::CreateFilter sizes (util/bloom.cc:30-38).
__global__ void k_window_sizes(const u32* nw_dev, EncParams ep, EncBufs b) {
  const u64 nw = *nw_dev;
  GRID_STRIDE(w, nw + 1) {
    const u64 n = w < nw ? b.w_keys[w] : 0;
    u64 sz = 0;
    if (n) {
      u64 bits = n * ep.bloom_bits;
      if (bits < 64) bits = 64;
      sz = (bits + 7) / 8 + 1;
    }
    b.w_fsz[w] = sz;
  }
}

__device__ __forceinline__ u64 meta_block_size(const EncParams& ep, u64 dend, u64 fb) {
  if (!ep.bloom_bits) return 8;
  const u64 hl = varint_len(dend) + varint_len(fb);
  return 3 + 34 + hl + 4 + 4;
}

__global__ void k_file_sizes(const u32* nf_dev, EncParams ep, EncBufs b) {
  const u64 nf = *nf_dev;
  GRID_STRIDE(f, nf + 1) {
    if (f == nf) {
      b.f_off[f] = 0;
      continue;
    }
    const u32 b0 = b.f_first[f], b1 = b.f_first[f + 1];
    const u64 dend = b.f_dend[f];
    u64 fb = 0, size = dend;
    if (ep.bloom_bits) {
      const u32 w0 = b.f_win[f], w1 = b.f_win[f + 1];
      const u64 fbytes = b.w_fsz[w1] - b.w_fsz[w0];
      b.f_fbytes[f] = fbytes;
      fb = fbytes + 4ull * (w1 - w0) + 5;
      size += fb + 5;
    }
    const u64 ib = b.ob_isz[b1] - b.ob_isz[b0] + 4ull * (b1 - b0) + 4;
    b.f_ibytes[f] = ib;
    size += meta_block_size(ep, dend, fb) + 5 + ib + 5 + 48;
    b.f_off[f] = size;
  }
}

// ---------------------------------------------------------------------------
// writers

__device__ __forceinline__ void load_crc_tables(u32 (*sl)[256], u32 (*z)[256]) {
  for (u32 i = threadIdx.x; i < 4 * 256; i += blockDim.x) {
    sl[i >> 8][i & 255] = g_crc.slice[i >> 8][i & 255];
    z[i >> 8][i & 255] = g_crc.zeros[6][i >> 8][i & 255];
  }
}


// The internal key user_key || fixed64(tag) of entry q as NWMAX big-endian
// words (bytes past the key are zero); the planes hold the user key's words.
template <u32 NWMAX>
__device__ __forceinline__ void ikey_words(const Planes& K, u64 q, u32 ulen, u64 tag, u64 (&ik)[NWMAX]) {
  const u64 tb = __builtin_bswap64(tag);
  const u32 wi = ulen >> 3, sft = (ulen & 7) * 8;
#pragma unroll
  for (u32 w = 0; w < NWMAX; w++) {
    u64 x = w < K.W ? K.P[(u64)w * K.stride + q] : 0;
    if (w == wi) x |= tb >> sft;
    if (sft && w == wi + 1) x |= tb << (64 - sft);
    ik[w] = x;
  }
}

// Byte j of big-endian words w.
template <u32 N>
__device__ __forceinline__ u32 word_byte(const u64 (&w)[N], u32 j) {
  u64 x = 0;
#pragma unroll
  for (u32 t = 0; t < N; t++)
    if (t == (j >> 3)) x = w[t];
  return (u32)(x >> (56 - 8 * (j & 7))) & 0xff;
}

// The index key after a block whose last entry is a, in NWMAX big-endian
// words, and its length: InternalKeyComparator::FindShortestSeparator
// against b's key (db/dbformat.cc:69-86, util/comparator.cc:29-48) or, with
// no b, FindShortSuccessor (db/dbformat.cc:88-101, util/comparator.cc:50-63).
// A shortened user key gets the tag PackSequenceAndType(kMaxSequenceNumber,
// kValueTypeForSeek); otherwise the key is a's internal key.
template <u32 NWMAX>
__device__ __forceinline__ u32 index_key_words(const Planes& K, u64 a, bool has_b, u64 bq, u64 (&out)[NWMAX]) {
  const u32 la = K.klen[a];
  u64 ua[NWMAX];
#pragma unroll
  for (u32 w = 0; w < NWMAX; w++) ua[w] = w < K.W ? K.P[(u64)w * K.stride + a] : 0;
  u32 cut = 0xffffffffu;  // user-key bytes kept before the incremented byte
  u32 inc = 0;            // the incremented byte
  if (has_b) {
    const u32 lb = K.klen[bq];
    const u32 m = min(la, lb);
    u32 d = m;
#pragma unroll
    for (u32 w = 0; w < NWMAX; w++) {
      const u64 x = w < K.W ? ua[w] ^ K.P[(u64)w * K.stride + bq] : 0;
      if (x && d == m) d = min(8 * w + ((u32)__clzll((long long)x) >> 3), m);
    }
    if (d < m) {
      const u32 da = word_byte(ua, d);
      u32 db = 0;
#pragma unroll
      for (u32 w = 0; w < NWMAX; w++)
        if (w == (d >> 3) && w < K.W) db = (u32)(K.P[(u64)w * K.stride + bq] >> (56 - 8 * (d & 7))) & 0xff;
      if (da < 0xff && da + 1 < db && d + 1 < la) {
        cut = d;
        inc = da + 1;
      }
    }
  } else {
    u32 i = la;
#pragma unroll
    for (u32 w = 0; w < NWMAX; w++) {
      const u64 nb = ~ua[w];  // first byte that is not 0xff
      if (nb && i == la) i = min(8 * w + ((u32)__clzll((long long)nb) >> 3), la);
    }
    if (i + 1 < la) {
      cut = i;
      inc = word_byte(ua, i) + 1;
    }
  }
  const bool shortened = cut != 0xffffffffu;
  const u32 plen = shortened ? cut : la;                 // user-key bytes copied
  const u32 at = shortened ? cut + 1 : la;                // where the tag goes
  const u64 tb = __builtin_bswap64(shortened ? kMaxSeqTag : K.tag[a]);
  const u32 wi = at >> 3, sft = (at & 7) * 8;
#pragma unroll
  for (u32 w = 0; w < NWMAX; w++) {
    const u32 nb = plen > 8 * w ? min(plen - 8 * w, 8u) : 0u;
    u64 x = nb ? ua[w] & (~0ull << (64 - 8 * nb)) : 0ull;
    if (shortened && w == (cut >> 3)) x |= (u64)inc << (56 - 8 * (cut & 7));
    if (w == wi) x |= tb >> sft;
    if (sft && w == wi + 1) x |= tb << (64 - sft);
    out[w] = x;
  }
  return at + 8;
}

// The bytes of output block i (entries, restart array, type byte, masked CRC)
// assembled at buf: BlockBuilder::Add/Finish (table/block_builder.cc:62-114)
// with the fork's (offset << 8) | count restart words, and
// TableBuilder::WriteRawBlock (table/table_builder.cc:178-195).  buf is
// either the wave's LDS stage or the block's place in the output; the two
// are separate instantiations so that every access is a plain LDS (ds_*) or
// global access and never a generic (flat) one that could address either.
// One lane per entry: header, key delta from the key words in registers,
// restart word, and the value copied with up to 16 loads in flight.
// NWMAX = 0 reads key bytes from the planes one at a time (any key width).
template <u32 NWMAX, typename Buf>
__device__ __forceinline__ void assemble_block(const Kept& k, const EncParams& ep, const EncBufs& b, u64 q0, u32 cnt,
                                               u64 bytes, Buf buf, const u32 (*sl)[256], const u32 (*z256)[256],
                                               DevStatus* st) {
  const u32 lane = threadIdx.x & 63;
  const u32 R = ep.interval;
  const Planes& K = k.K;
  const u32 ngroups = (cnt + R - 1) / R;
  const u64 dend = bytes - 4ull * ngroups - 4;  // where the restart array starts
  u64 base = 0;
  for (u32 j0 = 0; j0 < cnt; j0 += 64) {
    const u32 j = j0 + lane;
    const bool act = j < cnt;
    const u64 q = q0 + j;
    const bool rs = (j % R) == 0;
    const u32 sz = act ? (rs ? b.szr[q] : b.sznr[q]) : 0;
    const u32 incl = wave_incl_scan(sz);
    const u64 off = base + incl - sz;
    u64 my_voff = 0, my_vsrc = 0;
    u32 my_vl = 0;
    if (act) {
      const u32 ulen = K.klen[q];
      const u32 il = ulen + 8;
      const u32 sh = rs ? 0 : b.shared[q];
      const u32 vl = k.vlen[q];
      const GMEM u8* vs = gbytes(k.vptr[q]);
      const u64 tag = K.tag[q];
      u64 p = off;
      p = buf.put_varint(p, sh);
      p = buf.put_varint(p, il - sh);
      p = buf.put_varint(p, vl);
      if constexpr (NWMAX > 0) {
        u64 ik[NWMAX];
        ikey_words<NWMAX>(K, q, ulen, tag, ik);
#pragma unroll
        for (u32 w = 0; w < NWMAX; w++)
#pragma unroll
          for (u32 t = 0; t < 8; t++) {
            const u32 x = 8 * w + t;
            if (x >= sh && x < il) buf.put(p + (x - sh), (u8)(ik[w] >> (56 - 8 * t)));
          }
      } else {
        for (u32 t = sh; t < il; t++) buf.put(p + (t - sh), ikey_byte(K, q, ulen, tag, t));
      }
      p += il - sh;
      if (rs) buf.put32(dend + 4ull * (j / R), ((u32)off << 8) | min(R, cnt - j));
      if constexpr (!Buf::kLds) {
        for (u32 x = 0; x < vl; x += 64) {
          u32 v4[16];
#pragma unroll
          for (u32 t = 0; t < 16; t++)
            if (x + 4 * t + 4 <= vl) v4[t] = gload_u32(vs + x + 4 * t);
#pragma unroll
          for (u32 t = 0; t < 16; t++)
            if (x + 4 * t + 4 <= vl) buf.put32(p + x + 4 * t, v4[t]);
        }
        for (u32 x = vl & ~3u; x < vl; x++) buf.put(p + x, vs[x]);
      }
      my_voff = p;
      my_vl = vl;
      my_vsrc = k.vptr[q];
    }
    if constexpr (Buf::kLds) copy_values_chunked(buf, act, my_voff, my_vsrc, my_vl);
    base += wave_last(incl);
  }
  if (lane == 0) {
    buf.put32(dend + 4ull * ngroups, ngroups);
    if (base != dend) set_err(st, ERR_CORRUPT, SITE_CAPACITY);
    buf.put(bytes, 0);  // kNoCompression
  }
  wave_sync();
  const u32 crc = buf.crc(bytes + 1, sl, z256);
  if (lane == 0) buf.put32(bytes + 1, crc_mask(crc));
  wave_sync();
}

// Where the values of one group of up to 64 entries go: per entry lane, its
// first 16-byte chunk in the group's chunk sequence, its offset in the
// block, its length and its source address.
struct ValueTab {
  u32 cp[64];
  u32 off[64];
  u32 vl[64];
  u64 src[64];
};

// Byte sinks for assemble_block.  LdsBuf points into the wave's __shared__
// stage (the block starts `mis` bytes in) and nothing else, so the compiler
// addresses it with ds_* instructions.
struct LdsBuf {
  static constexpr bool kLds = true;
  u8* base;  // the stage, 16-byte aligned
  u32 mis;
  ValueTab* vt;
  __device__ __forceinline__ void put(u64 x, u8 v) const { base[mis + x] = v; }
  __device__ __forceinline__ u8 get(u64 x) const { return base[mis + x]; }
  __device__ __forceinline__ void put32(u64 x, u32 v) const {
    u8* p = base + mis + x;
    p[0] = (u8)v;
    p[1] = (u8)(v >> 8);
    p[2] = (u8)(v >> 16);
    p[3] = (u8)(v >> 24);
  }
  __device__ __forceinline__ u64 put_varint(u64 x, u64 v) const {
    u8* p = base + mis;
    while (v >= 128) {
      p[x++] = (u8)(v | 128);
      v >>= 7;
    }
    p[x++] = (u8)v;
    return x;
  }
  __device__ __forceinline__ u32 crc(u64 L, const u32 (*sl)[256], const u32 (*z256)[256]) const {
    const u64 rows = (L + 255) >> 8;
    return wave_crc32c_rows_lds(L, 0, rows, reinterpret_cast<const u32*>(base), mis, sl, z256) ^ 0xffffffffu;
  }
};
struct GlobalBuf {
  static constexpr bool kLds = false;
  u8* p;
  __device__ __forceinline__ void put(u64 x, u8 v) const { p[x] = v; }
  __device__ __forceinline__ u8 get(u64 x) const { return p[x]; }
  __device__ __forceinline__ void put32(u64 x, u32 v) const {
    p[x] = (u8)v;
    p[x + 1] = (u8)(v >> 8);
    p[x + 2] = (u8)(v >> 16);
    p[x + 3] = (u8)(v >> 24);
  }
  __device__ __forceinline__ u64 put_varint(u64 x, u64 v) const {
    while (v >= 128) {
      p[x++] = (u8)(v | 128);
      v >>= 7;
    }
    p[x++] = (u8)v;
    return x;
  }
  __device__ __forceinline__ u32 crc(u64 L, const u32 (*sl)[256], const u32 (*z256)[256]) const {
    return wave_crc32c(L, sl, z256, [&](u64 x) -> u8 { return p[x]; });
  }
};

// The values of a group of entries (lanes with act) copied into an LDS
// block as 16-byte chunks spread over the whole wave: lane c of a round
// takes chunk c of the group, found by a binary search of the chunk prefix.
// Sources are read 16 bytes at a time; they are always followed by at least
// 15 more readable bytes (a table's restart array, trailer and footer, or
// the padded value buffer of luda_build_tables).
__device__ __forceinline__ void copy_values_chunked(const LdsBuf& buf, bool act, u64 voff, u64 vsrc, u32 vl) {
  const u32 lane = threadIdx.x & 63;
  ValueTab& vt = *buf.vt;
  const u32 nch = act ? (vl + 15) >> 4 : 0;
  const u32 cincl = wave_incl_scan(nch);
  const u32 C = wave_last(cincl);
  vt.cp[lane] = cincl - nch;
  vt.off[lane] = (u32)voff;
  vt.vl[lane] = vl;
  vt.src[lane] = vsrc;
  wave_sync();
  constexpr u32 kRounds = 4;
  for (u32 c0 = 0; c0 < C; c0 += 64 * kRounds) {
    u32 d[kRounds][4];
    u32 dst[kRounds], n[kRounds];
#pragma unroll
    for (u32 r = 0; r < kRounds; r++) {
      const u32 c = c0 + 64 * r + lane;
      n[r] = 0;
      dst[r] = 0;
      if (c < C) {
        u32 lo = 0, hi = 64;  // last entry with cp <= c
        while (hi - lo > 1) {
          const u32 mid = (lo + hi) >> 1;
          if (vt.cp[mid] <= c) lo = mid;
          else hi = mid;
        }
        const u32 k = c - vt.cp[lo];
        const GMEM u8* src = gbytes(vt.src[lo] + 16ull * k);
        n[r] = min(16u, vt.vl[lo] - 16 * k);
        dst[r] = vt.off[lo] + 16 * k;
#pragma unroll
        for (u32 t = 0; t < 4; t++) d[r][t] = gload_u32(src + 4 * t);
      }
    }
#pragma unroll
    for (u32 r = 0; r < kRounds; r++) {
#pragma unroll
      for (u32 t = 0; t < 4; t++) {
        if (4 * t + 4 <= n[r]) {
          buf.put32(dst[r] + 4 * t, d[r][t]);
        } else {
#pragma unroll
          for (u32 b = 0; b < 4; b++)
            if (4 * t + b < n[r]) buf.put(dst[r] + 4 * t + b, (u8)(d[r][t] >> (8 * b)));
        }
      }
    }
  }
  wave_sync();
}

// Copy n bytes from the LDS stage (src, which agrees with dst mod 16) to dst.
__device__ __forceinline__ void wave_copy_out(u8* dst, const uint4* stage4, u32 mis, u64 n) {
  const u32 lane = threadIdx.x & 63;
  const u8* stage = reinterpret_cast<const u8*>(stage4);
  const u64 head = (16 - mis) & 15;
  const u64 h = head < n ? head : n;
  if (lane < h) dst[lane] = stage[mis + lane];
  const u64 body = (n - h) & ~15ull;
  // stage[mis + h] is 16-byte aligned: mis + h is 0 or 16
  const u32 s4 = (u32)((mis + h) >> 4);
  uint4* d4 = reinterpret_cast<uint4*>(dst + h);
  for (u64 x = lane; x < body / 16; x += 64) d4[x] = stage4[s4 + x];
  const u64 t0 = h + body;
  if (t0 + lane < n) dst[t0 + lane] = stage[mis + t0 + lane];
}

// Where output block i goes and which entries it holds.
struct BlockMeta {
  u64 off;    // offset in its table
  u32 file;   // its table
  u32 first;  // first entry
  u32 n;      // entries
  u32 bytes;  // block bytes without the trailer
};
__device__ __forceinline__ BlockMeta block_meta(const EncBufs& b, u64 i) {
  BlockMeta m;
  m.off = b.ob_off[i];
  m.file = b.ob_file[i];
  m.first = b.ob_first[i];
  m.n = b.ob_n[i];
  m.bytes = b.ob_bytes[i];
  return m;
}

// Each wavefront writes data blocks i, i + S, i + 2S, ... (S = all waves of
// the grid).  Blocks that fit are assembled in the wave's LDS stage and
// written out with 16-byte stores; larger ones (a single big value) are
// assembled in place.  The block after next has its metadata, and the next
// its table offset, loaded while the current one is assembled.
template <u32 NWMAX>
__global__ __launch_bounds__(kThreads) void k_write_blocks(Kept k, const u32* nbo_dev, EncParams ep, EncBufs b,
                                                            u8* out, u64 out_cap, DevStatus* st) {
  __shared__ u32 sl[4][256];
  __shared__ u32 z256[4][256];
  __shared__ uint4 stage_all[kWaves][(kStage + 32) / 16];
  __shared__ ValueTab vtab[kWaves];
  if (st->out_bytes > out_cap || st->err) return;
  load_crc_tables(sl, z256);
  __syncthreads();
  const u64 nbo = *nbo_dev;
  const u32 wave = threadIdx.x >> 6;
  const u64 S = (u64)gridDim.x * kWaves;
  u64 i = (u64)blockIdx.x * kWaves + wave;
  BlockMeta m0 = {0, 0, 0, 0, 0}, m1 = {0, 0, 0, 0, 0};
  u64 d0 = 0;
  if (i < nbo) {
    m0 = block_meta(b, i);
    d0 = b.f_off[m0.file] + m0.off;
  }
  if (i + S < nbo) m1 = block_meta(b, i + S);
  for (; i < nbo; i += S) {
    u64 d1 = 0;
    BlockMeta m2 = {0, 0, 0, 0, 0};
    if (i + S < nbo) d1 = b.f_off[m1.file] + m1.off;
    if (i + 2 * S < nbo) m2 = block_meta(b, i + 2 * S);
    u8* dst = out + d0;
    const u64 bytes = m0.bytes;
    if (bytes + 5 + 16 <= kStage) {
      const u32 mis = (u32)((u64)dst & 15);
      assemble_block<NWMAX>(k, ep, b, m0.first, m0.n, bytes,
                            LdsBuf{reinterpret_cast<u8*>(stage_all[wave]), mis, &vtab[wave]}, sl, z256, st);
      wave_copy_out(dst, stage_all[wave], mis, bytes + 5);
    } else {
      assemble_block<NWMAX>(k, ep, b, m0.first, m0.n, bytes, GlobalBuf{dst}, sl, z256, st);
    }
    wave_sync();
    m0 = m1;
    d0 = d1;
    m1 = m2;
  }
}

// One wavefront per filter window with keys: bloom bits with LDS atomics.
__global__ __launch_bounds__(kThreads) void k_write_filters(Kept k, const u32* nw_dev, const u32* nf_dev,
                                                             EncParams ep, EncBufs b, u8* out, u64 out_cap,
                                                             DevStatus* st) {
  __shared__ u32 bits_all[kWaves][kFBits / 4];
  if (st->out_bytes > out_cap || st->err) return;
  const u64 nw = *nw_dev;
  const u32 nf = *nf_dev;
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  u32* lb = bits_all[wave];
  for (u64 w = (u64)blockIdx.x * kWaves + wave; w < nw; w += (u64)gridDim.x * kWaves) {
    const u32 n = b.w_keys[w];
    if (!n) continue;
    // table of this window
    u32 lo = 0, hi = nf;
    while (hi - lo > 1) {
      const u32 mid = (lo + hi) >> 1;
      if (b.f_win[mid] <= w) lo = mid;
      else hi = mid;
    }
    const u32 f = lo;
    u8* dst = out + b.f_off[f] + b.f_dend[f] + (b.w_fsz[w] - b.w_fsz[b.f_win[f]]);
    u64 nbits = (u64)n * ep.bloom_bits;
    if (nbits < 64) nbits = 64;
    const u64 nbytes = (nbits + 7) / 8;
    nbits = nbytes * 8;
    if (nbytes > kFBits) {
      if (lane == 0) set_err(st, ERR_UNSUPPORTED, SITE_CAPACITY);
      continue;
    }
    for (u32 i = lane; i < (nbytes + 3) / 4; i += 64) lb[i] = 0;
    wave_sync();
    const u64 q0 = b.w_first[w];
    for (u32 t = lane; t < n; t += 64) {
      u32 h = hash_ukey(k.K, q0 + t, 0xbc9f1d34u);
      const u32 delta = (h >> 17) | (h << 15);
      for (u32 j = 0; j < ep.bloom_k; j++) {
        const u32 bp = (u32)(h % nbits);
        atomicOr(&lb[bp >> 5], 1u << (bp & 31));
        h += delta;
      }
    }
    wave_sync();
    for (u32 i = lane; i < nbytes; i += 64) dst[i] = (u8)(lb[i >> 2] >> (8 * (i & 3)));
    if (lane == 0) dst[nbytes] = (u8)ep.bloom_k;
    wave_sync();
  }
}

__device__ __forceinline__ u8* put_fixed32(u8* p, u32 v) {
  st32(p, v);
  return p + 4;
}

// crc32c::Value of L bytes with the rows split over the workgroup's waves:
// each wave folds a contiguous run of rows, and the runs are joined by
// appending to each the zero bytes of the rows after it.  Called by every
// thread of the workgroup; every thread gets the result.
template <typename ByteAt>
__device__ __forceinline__ u32 block_crc_split(u64 L, const u32 (*sl)[256], const u32 (*z256)[256], u32* part,
                                               ByteAt at) {
  const u32 wave = threadIdx.x >> 6;
  const u64 rows = (L + 255) >> 8;
  const u64 per = (rows + kWaves - 1) / kWaves;
  const u64 r0 = min(rows, per * wave), r1 = min(rows, r0 + per);
  u32 s = wave_crc32c_rows(L, r0, r1, sl, z256, at);
  s = crc_append_zeros(s, 256 * (rows - r1));
  if ((threadIdx.x & 63) == 0) part[wave] = s;
  __syncthreads();
  u32 x = 0;
#pragma unroll
  for (u32 w = 0; w < kWaves; w++) x ^= part[w];
  __syncthreads();
  return x ^ 0xffffffffu;
}

// One workgroup per table: filter offsets, filter/metaindex/index blocks with
// trailers, footer, and the table's metadata.
template <u32 NWMAX>
__global__ __launch_bounds__(kThreads) void k_write_tail(Kept k, const u32* nf_dev, EncParams ep, EncBufs b, u8* out,
                                                          u64 out_cap, DevStatus* st) {
  __shared__ u32 sl[4][256];
  __shared__ u32 z256[4][256];
  __shared__ u64 s_moff, s_msize;
  __shared__ u32 s_part[kWaves];
  if (st->out_bytes > out_cap || st->err) return;
  load_crc_tables(sl, z256);
  __syncthreads();
  const u64 nf = *nf_dev;
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const Planes& K = k.K;
  for (u64 f = blockIdx.x; f < nf; f += gridDim.x) {
    const u32 b0 = b.f_first[f], b1 = b.f_first[f + 1];
    u8* base = out + b.f_off[f];
    const u64 dend = b.f_dend[f];
    u64 fbsize = 0;
    // filter block: offsets array, array offset, base lg (FilterBlockBuilder::Finish)
    if (ep.bloom_bits) {
      const u32 w0 = b.f_win[f], w1 = b.f_win[f + 1];
      const u64 fbytes = b.f_fbytes[f];
      u8* fs = base + dend;
      for (u32 i = threadIdx.x; i < w1 - w0; i += blockDim.x)
        st32(fs + fbytes + 4ull * i, (u32)(b.w_fsz[w0 + i] - b.w_fsz[w0]));
      if (threadIdx.x == 0) {
        st32(fs + fbytes + 4ull * (w1 - w0), (u32)fbytes);
        fs[fbytes + 4ull * (w1 - w0) + 4] = 11;
      }
      fbsize = fbytes + 4ull * (w1 - w0) + 5;
    }
    const u64 moff = dend + (ep.bloom_bits ? fbsize + 5 : 0);
    const u64 msize = meta_block_size(ep, dend, fbsize);
    const u64 ioff = moff + msize + 5;
    const u64 isize = b.f_ibytes[f];
    // metaindex block (one BlockBuilder entry, table_builder.cc:216-227)
    if (threadIdx.x == 0) {
      u8* p = base + moff;
      if (ep.bloom_bits) {
        static const char kName[] = "filter.leveldb.BuiltinBloomFilter2";
        const u32 hl = varint_len(dend) + varint_len(fbsize);
        *p++ = 0;
        *p++ = 34;
        *p++ = (u8)hl;
        for (u32 i = 0; i < 34; i++) *p++ = (u8)kName[i];
        p = enc_varint(enc_varint(p, dend), fbsize);
        p = put_fixed32(p, 1);  // (0 << 8) | 1
        p = put_fixed32(p, 1);
      } else {
        p = put_fixed32(p, 0);  // (0 << 8) | 0
        p = put_fixed32(p, 1);
      }
      s_moff = moff;
      s_msize = msize;
    }
    // index block entries, restart interval 1
    {
      u8* ib = base + ioff;
      const u64 ibase = b.ob_isz[b0];
      const u64 isum = b.ob_isz[b1] - ibase;
      for (u32 i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
        const u64 eoff = b.ob_isz[i] - ibase;
        u8* p = ib + eoff;
        const u64 last = (u64)b.ob_first[i] + b.ob_n[i] - 1;
        const bool has_b = i + 1 < b1;
        const u64 boff = b.ob_off[i], bsz = b.ob_bytes[i];
        const u32 hl = varint_len(boff) + varint_len(bsz);
        if constexpr (NWMAX > 0) {
          // the key in registers first, then its bytes: no load waits on a store
          u64 kw[NWMAX];
          const u32 kl = index_key_words<NWMAX>(K, last, has_b, has_b ? b.ob_first[i + 1] : 0, kw);
          *p++ = 0;
          p = enc_varint(p, kl);
          p = enc_varint(p, hl);
#pragma unroll
          for (u32 w = 0; w < NWMAX; w++)
#pragma unroll
            for (u32 t = 0; t < 8; t++)
              if (8 * w + t < kl) p[8 * w + t] = (u8)(kw[w] >> (56 - 8 * t));
          p += kl;
        } else {
          const u32 kl = has_b ? index_sep(K, last, b.ob_first[i + 1], nullptr) : index_succ(K, last, nullptr);
          *p++ = 0;
          p = enc_varint(p, kl);
          p = enc_varint(p, hl);
          if (has_b) index_sep(K, last, b.ob_first[i + 1], p);
          else index_succ(K, last, p);
          p += kl;
        }
        p = enc_varint(p, boff);
        p = enc_varint(p, bsz);
        st32(ib + isum + 4ull * (i - b0), ((u32)eoff << 8) | 1u);
      }
      if (threadIdx.x == 0) st32(ib + isum + 4ull * (b1 - b0), b1 - b0);
    }
    // type bytes, then the filter and index blocks' CRCs with the rows of
    // each split over the four waves (the long ones: ~47 KB of filters per
    // 2 MiB table), the metaindex's by wave 0
    if (threadIdx.x == 0) {
      if (ep.bloom_bits) base[dend + fbsize] = 0;
      base[ioff + isize] = 0;
      base[s_moff + s_msize] = 0;
    }
    __syncthreads();
    if (ep.bloom_bits) {
      u8* fs = base + dend;
      const u32 crc = block_crc_split(fbsize + 1, sl, z256, s_part, [&](u64 x) -> u8 { return fs[x]; });
      if (threadIdx.x == 0) st32(fs + fbsize + 1, crc_mask(crc));
    }
    {
      u8* is = base + ioff;
      const u32 crc = block_crc_split(isize + 1, sl, z256, s_part, [&](u64 x) -> u8 { return is[x]; });
      if (threadIdx.x == 0) st32(is + isize + 1, crc_mask(crc));
    }
    if (wave == 0) {
      u8* ms = base + s_moff;
      const u32 crc = wave_crc32c(s_msize + 1, sl, z256, [&](u64 x) -> u8 { return ms[x]; });
      if (lane == 0) st32(ms + s_msize + 1, crc_mask(crc));
    } else if (wave == 3 && lane == 0) {
      // footer (Footer::EncodeTo, table/format.cc:31-40)
      u8* p = base + ioff + isize + 5;
      u8* s = p;
      p = enc_varint(p, moff);
      p = enc_varint(p, msize);
      p = enc_varint(p, ioff);
      p = enc_varint(p, isize);
      while (p < s + 40) *p++ = 0;
      p = put_fixed32(p, 0x8b80fb57u);
      p = put_fixed32(p, 0xdb477524u);
      // table metadata: Output::smallest / largest (db/db_impl.cc:1163-1166)
      const u64 qs = b.ob_first[b0];
      const u64 ql = (u64)b.ob_first[b1 - 1] + b.ob_n[b1 - 1] - 1;
      FileMeta fm;
      fm.offset = b.f_off[f];
      fm.size = (u64)(p - base);
      fm.nent = ql + 1 - qs;
      fm.slen = K.klen[qs] + 8;
      fm.llen = K.klen[ql] + 8;
      b.meta[f] = fm;
      u8* ks = b.keys + f * 2ull * ep.kw;
      for (u32 t = 0; t < fm.slen; t++) ks[t] = ikey_byte(K, qs, fm.slen - 8, K.tag[qs], t);
      u8* kl = ks + ep.kw;
      for (u32 t = 0; t < fm.llen; t++) kl[t] = ikey_byte(K, ql, fm.llen - 8, K.tag[ql], t);
    }
    __syncthreads();
  }
}

u32 grid_for(u64 n, u32 per_block) {
  u64 g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return (u32)g;
}

}  // namespace

void launch_entry_sizes(KTimer& T, Kept k, const u32* m_dev, u32 m_cap, EncBufs b, DevStatus* st) {
  KLAUNCH(T, "entry_sizes", k_entry_sizes, grid_for(m_cap, 256), 256, 0, k, m_dev, b, st);
}
void launch_bnext(KTimer& T, Kept k, const u32* m_dev, u32 m_cap, EncParams ep, EncBufs b, DevStatus* st) {
  KLAUNCH(T, "bnext", k_bnext, std::min<u32>(grid_for(m_cap, 256), 2048), 256, 0, k, m_dev, ep, b, st);
}
void launch_block_info(KTimer& T, Kept k, const u32* m_dev, u32 m_cap, EncBufs b, DevStatus* st) {
  KLAUNCH(T, "block_info", k_block_info, grid_for(m_cap, 256), 256, 0, k, m_dev, b, st);
}
void launch_block_tot(KTimer& T, const u32* nbo_dev, u32 nbo_cap, EncBufs b, u32* tot) {
  KLAUNCH(T, "block_tot", k_block_tot, grid_for(nbo_cap + 1, 256), 256, 0, nbo_dev, b, tot);
}
void launch_block_forced(KTimer& T, const u32* nbo_dev, u32 nbo_cap, EncBufs b) {
  KLAUNCH(T, "block_forced", k_block_forced, grid_for(nbo_cap, 256), 256, 0, nbo_dev, b);
}
void launch_fnext(KTimer& T, const u32* nbo_dev, u32 nbo_cap, const u32* nforced_dev, EncParams ep, EncBufs b,
                  DevStatus* st) {
  KLAUNCH(T, "fnext", k_fnext, grid_for(nbo_cap, 256), 256, 0, nbo_dev, nforced_dev, ep, b, st);
}
void launch_file_first(KTimer& T, const u32* nbo_dev, u32 nbo_cap, EncBufs b, const u32* nf_dev) {
  KLAUNCH(T, "file_first", k_file_first, grid_for(nbo_cap, 256), 256, 0, nbo_dev, b, nf_dev);
}
void launch_block_layout(KTimer& T, Kept k, const u32* nbo_dev, u32 nbo_cap, const u32* nf_dev, EncParams ep,
                         EncBufs b) {
  KLAUNCH(T, "isz_sentinel", k_isz_sentinel, 1, 1, 0, nbo_dev, b);
  KLAUNCH(T, "block_layout", k_block_layout, grid_for(nbo_cap, 256), 256, 0, k, nbo_dev, nf_dev, ep, b);
}
void launch_file_layout(KTimer& T, const u32* nf_dev, u32 nf_cap, EncParams ep, EncBufs b) {
  KLAUNCH(T, "file_layout", k_file_layout, grid_for(nf_cap, 256), 256, 0, nf_dev, ep, b);
}
void launch_window_init(KTimer& T, const u32* nw_dev, u32 nw_cap, EncBufs b) {
  KLAUNCH(T, "window_init", k_window_init, grid_for(nw_cap + 1, 256), 256, 0, nw_dev, b);
}
void launch_window_counts(KTimer& T, const u32* nbo_dev, u32 nbo_cap, const u32* nf_dev, const u32* nw_dev,
                          EncBufs b) {
  KLAUNCH(T, "window_counts", k_window_counts, grid_for(nbo_cap, 256), 256, 0, nbo_dev, nf_dev, nw_dev, b);
}
void launch_window_sizes(KTimer& T, const u32* nw_dev, u32 nw_cap, EncParams ep, EncBufs b) {
  KLAUNCH(T, "window_sizes", k_window_sizes, grid_for(nw_cap + 1, 256), 256, 0, nw_dev, ep, b);
}
void launch_file_sizes(KTimer& T, const u32* nf_dev, u32 nf_cap, EncParams ep, EncBufs b) {
  KLAUNCH(T, "file_sizes", k_file_sizes, grid_for(nf_cap + 1, 256), 256, 0, nf_dev, ep, b);
}
void launch_write_blocks(KTimer& T, Kept k, const u32* nbo_dev, u32 nbo_cap, EncParams ep, EncBufs b, u8* out,
                         u64 out_cap, DevStatus* st) {
  // enough workgroups to fill every CU a few times; each wave loops
  const u32 g = std::min<u32>(grid_for(nbo_cap, kWaves), 2048);
  if (k.K.W <= 2)
    KLAUNCH(T, "write_blocks", k_write_blocks<3>, g, kThreads, 0, k, nbo_dev, ep, b, out, out_cap, st);
  else if (k.K.W <= 4)
    KLAUNCH(T, "write_blocks", k_write_blocks<5>, g, kThreads, 0, k, nbo_dev, ep, b, out, out_cap, st);
  else if (k.K.W <= 8)
    KLAUNCH(T, "write_blocks", k_write_blocks<9>, g, kThreads, 0, k, nbo_dev, ep, b, out, out_cap, st);
  else
    KLAUNCH(T, "write_blocks", k_write_blocks<0>, g, kThreads, 0, k, nbo_dev, ep, b, out, out_cap, st);
}
void launch_write_filters(KTimer& T, Kept k, const u32* nw_dev, u32 nw_cap, const u32* nf_dev, EncParams ep, EncBufs b,
                          u8* out, u64 out_cap, DevStatus* st) {
  KLAUNCH(T, "write_filters", k_write_filters, std::min<u32>(grid_for(nw_cap, kWaves), 2048), kThreads, 0, k, nw_dev,
          nf_dev, ep, b, out,
          out_cap, st);
}
void launch_write_tail(KTimer& T, Kept k, const u32* nf_dev, u32 nf_cap, EncParams ep, EncBufs b, u8* out,
                       u64 out_cap, DevStatus* st) {
  const u32 g = grid_for(nf_cap, 1);
  if (k.K.W <= 2)
    KLAUNCH(T, "write_tail", k_write_tail<3>, g, kThreads, 0, k, nf_dev, ep, b, out, out_cap, st);
  else if (k.K.W <= 4)
    KLAUNCH(T, "write_tail", k_write_tail<5>, g, kThreads, 0, k, nf_dev, ep, b, out, out_cap, st);
  else if (k.K.W <= 8)
    KLAUNCH(T, "write_tail", k_write_tail<9>, g, kThreads, 0, k, nf_dev, ep, b, out, out_cap, st);
  else
    KLAUNCH(T, "write_tail", k_write_tail<0>, g, kThreads, 0, k, nf_dev, ep, b, out, out_cap, st);
}

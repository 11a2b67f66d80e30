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
#include "stage_values.h"
#include "scan.h"

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
// a table cut is due, or the stream ends.  With entries [q, y) in a block
// starting at q, the estimate is
//   S(q, y) = sum of sznr over [q, y) + sum of (szr - sznr) over its restart
//             points q, q + R, ... + 4 * (restart points) + 4,
// which only grows with y.  A workgroup takes 2048 starts (8 consecutive ones
// per thread: the whole grid of a 256 MiB job then fits the device at once,
// five workgroups per CU) and stages kBnSpan entries past them in LDS as a
// prefix sum of sznr (with the count of forced entries packed above it) and
// the restart surcharges.  Per start, the prefix alone bounds the cut from above
// (S >= prefix difference + 8); a backward walk then finds the exact first y
// with S >= block_size.  The bound only moves forward from one start to the
// next.  A block that runs past the staged span is finished by walking the
// entries in global memory.
constexpr u32 kBnPerThread = 8;                 // starts per thread
constexpr u32 kBnStarts = 256 * kBnPerThread;
constexpr u32 kBnSpan = 512;
constexpr u32 kBnTile = kBnStarts + kBnSpan;     // staged entries
constexpr u32 kBnPer = (kBnTile + 255) / 256;    // staged entries per thread
constexpr u64 kCfOne = 1ull << 48;               // one forced entry in the packed prefix
constexpr u64 kByteMask = kCfOne - 1;
__global__ __launch_bounds__(256) void k_bnext(Kept k, const u32* m_dev, EncParams ep, EncBufs b, DevStatus* st) {
  __shared__ u64 sP[kBnTile + 1];  // sP[x]: sznr of staged entries [0, x), plus kCfOne per forced one
  __shared__ u32 sD[kBnTile];      // szr - sznr: what a restart point adds (>= 0)
  __shared__ u64 s_wsum[4];
  const u64 m = *m_dev;
  const u32 R = ep.interval;
  const u64 bs = ep.block_size;
  const u32 tid = threadIdx.x, wave = tid >> 6;
  u32 smax = 0;
  for (u64 q0 = (u64)blockIdx.x * kBnStarts; q0 < m; q0 += (u64)gridDim.x * kBnStarts) {
    const u32 lim = (u32)min((u64)kBnTile, m - q0);  // staged entries [q0, q0 + lim)
    // stage: thread t owns staged entries [kBnPer t, kBnPer t + kBnPer)
    // coalesced loads (all issued before any is used) into LDS: the
    // surcharges to sD, the packed sizes to sP (shifted by one); then each
    // thread takes its kBnPer consecutive entries from there
    {
      u32 r[kBnPer], nr[kBnPer], fo[kBnPer];
#pragma unroll
      for (u32 i = 0; i < kBnPer; i++) {
        const u32 x = 256 * i + tid;
        const u64 qx = q0 + (x < lim ? x : 0);
        r[i] = b.szr[qx];
        nr[i] = b.sznr[qx];
        fo[i] = k.forced[qx];
      }
      bool bad = false;
#pragma unroll
      for (u32 i = 0; i < kBnPer; i++) {
        const u32 x = 256 * i + tid;
        if (x < lim) {
          bad |= r[i] >= (1u << 31) || nr[i] >= (1u << 31) || r[i] < nr[i];
          sD[x] = r[i] - nr[i];
          sP[x + 1] = (u64)nr[i] + (fo[i] ? kCfOne : 0);
        }
      }
      if (bad) set_err(st, ERR_UNSUPPORTED, SITE_BLOCK_ENTRIES);
    }
    __syncthreads();
    u64 v[kBnPer];
    u64 tot = 0;
#pragma unroll
    for (u32 i = 0; i < kBnPer; i++) {
      const u32 x = kBnPer * tid + i;
      v[i] = x < lim ? sP[x + 1] : 0;
      tot += v[i];
    }
    const u64 incl = wave_incl_sum(tot);
    if ((tid & 63) == 63) s_wsum[wave] = incl;
    __syncthreads();  // also: every thread has read its sizes before they are overwritten
    u64 run = incl - tot;
    for (u32 w = 0; w < wave; w++) run += s_wsum[w];
    if (tid == 0) sP[0] = 0;
#pragma unroll
    for (u32 i = 0; i < kBnPer; i++) {
      const u32 x = kBnPer * tid + i;
      run += v[i];
      if (x < lim) sP[x + 1] = run;
    }
    __syncthreads();
    u32 ub = 0;  // the upper bound of the previous start
    // entries per block at the tile's mean entry size: the first search's guess
    const u64 tbytes = sP[lim] & kByteMask;
    const u32 est = tbytes ? (u32)min<u64>((u64)lim, (bs * lim) / tbytes) : 1;
    for (u32 s = 0; s < kBnPerThread; s++) {
      const u32 x = kBnPerThread * tid + s;
      const u64 q = q0 + x;
      if (x >= kBnStarts || q >= m) break;
      const u64 px = sP[x];
      const u64 bx = px & kByteMask;
      // smallest y in (x, lim] with bytes(x, y) + 8 >= bs, else lim + 1:
      // for the thread's first start from the tile's mean entry size, for
      // the others from the previous start's bound, stepping at most 16
      // entries before a binary search takes over
      auto reaches = [&](u32 yy) { return (sP[yy] & kByteMask) - bx + 8 >= bs; };
      u32 y = s == 0 ? min(lim, x + 1 + est) : max(ub, x + 1);
      u32 lo = x + 1, hi = lim + 1;  // the answer lies in [lo, hi]
      bool done = false;
      for (u32 step = 0; step < 16 && !done; step++) {
        if (y > lim || reaches(y)) {
          hi = min(hi, y);
          if (y == lo || (y <= lim && !reaches(y - 1))) done = true;
          else y--;
        } else {
          lo = y + 1;
          y++;
          if (y > lim) {
            y = lim + 1;
            done = true;
          }
        }
      }
      if (!done) {
        while (lo < hi) {
          const u32 mid = (lo + hi) >> 1;
          if (reaches(mid)) hi = mid;
          else lo = mid + 1;
        }
        y = lo;
      }
      ub = y;
      // exact S(x, y0) for y0 = min(y, lim); g restart points
      const u32 y0 = min(y, lim);
      const u32 g = (y0 - x + R - 1) / R;
      u64 sur = 0;
      for (u32 t = 0; t < g; t++) sur += sD[x + t * R];
      u64 S = (sP[y0] & kByteMask) - bx + sur + 4ull * g + 4;
      u32 end;       // the cut, when it lies within the staged entries
      bool found = S >= bs;
      if (found) {
        // walk back while the block one entry shorter still reaches bs
        end = y0;
        while (end > x + 1) {
          const u32 e = end - 1;  // the entry that would be dropped
          const bool rs = (e - x) % R == 0;
          const u64 sz = (sP[e + 1] & kByteMask) - (sP[e] & kByteMask) + (rs ? (u64)sD[e] + 4 : 0);
          if (S - sz < bs) break;
          S -= sz;
          end = e;
        }
      } else {
        end = lim;  // not within the staged entries
      }
      // a forced entry in (x, end) cuts first
      const u64 cf0 = sP[x + 1] >> 48;
      if (end > x + 1 && (sP[end] >> 48) > cf0) {
        u32 lo = x + 1, hi = end - 1;  // smallest e in [x + 1, end) with forced[e]: sP[e + 1] counts it
        while (lo < hi) {
          const u32 mid = (lo + hi) >> 1;
          if ((sP[mid + 1] >> 48) > cf0) hi = mid;
          else lo = mid + 1;
        }
        end = lo;
        const u32 ge = (end - x + R - 1) / R;
        u64 se = 0;
        for (u32 t = 0; t < ge; t++) se += sD[x + t * R];
        S = (sP[end] & kByteMask) - bx + se + 4ull * ge + 4;
        found = true;
      }
      u64 j = q0 + end;
      if (!found && j < m) {
        // past the staged entries: carry on from global memory
        const u32 cnt = end - x;
        u64 acc = S - 4ull * g - 4;
        u32 ng = g, rpos = cnt % R;
        bool done = k.forced[j] != 0;
        while (!done) {
          const bool rs = rpos == 0;
          acc += rs ? b.szr[j] : b.sznr[j];
          ng += rs;
          rpos = rpos + 1 == R ? 0 : rpos + 1;
          j++;
          done = acc + 4ull * ng + 4 >= bs || j >= m || k.forced[j] != 0;
        }
        S = acc + 4ull * ng + 4;
      }
      b.bnext[q] = (u32)j;
      b.bbytes[q] = (u32)S;
      smax = max(smax, (u32)(j - q));
    }
    // the LDS is reused by the next tile (a barrier waits for this tile's
    // stores too, so it is skipped after the last one)
    if (q0 + (u64)gridDim.x * kBnStarts < m) __syncthreads();
  }
  smax = wave_max(smax);
  if ((threadIdx.x & 63) == 0 && smax) atomicMax(&st->blk_step_max, smax);
}

// The block scan's epilogue (over the block-start flags bs): output block id
// = the exclusive count of block starts before entry q.
struct BlockInfoEpi {
  const u32* forced;
  EncBufs b;
  u32 cap;  // output blocks the arrays hold
  struct Pre {
    u32 next, bytes, forced;
  };
  __device__ Pre pre(u64 q, u32 start) const {
    if (!start) return {0, 0, 0};
    return {b.bnext[q], b.bbytes[q], forced[q]};
  }
  __device__ void item(u64 q, u32 start, u32 id, const Pre& p) const {
    if (!start || id >= cap) return;
    b.ob_first[id] = (u32)q;
    b.ob_n[id] = p.next - (u32)q;
    b.ob_bytes[id] = p.bytes;
    b.ob_forced[id] = (q > 0 && p.forced) ? 1u : 0u;
  }
  __device__ void done(u32 nbo, u64) const {
    b.ob_forced[nbo] = 0;
    b.ob_bytes[nbo] = 0;
  }
};

// BloomFilterPolicy::CreateFilter sizes (util/bloom.cc:30-38).
__device__ __forceinline__ u64 filter_bytes(u64 keys, u32 bits_per_key) {
  if (!keys) return 0;
  u64 bits = keys * bits_per_key;
  if (bits < 64) bits = 64;
  return (bits + 7) / 8 + 1;
}

__device__ __forceinline__ u64 meta_block_size(const EncParams& ep, u64 dend, u64 fb) {
  if (!ep.bloom_bits) return 8;
  const u64 hl = varint_len(dend) + varint_len(fb);
  return 3 + 34 + hl + 4 + 4;
}

// ---------------------------------------------------------------------------
// Table cuts and table layout, over the output blocks.
//
// A table that starts at block t ends before the first block j > t with
// O[j] - O[t] >= max_file_size, O being the blocks' offsets in the output's
// data (FileSize() reached the limit after block j - 1,
// db/db_impl.cc:1173-1180), or that ShouldStopBefore starts (:1105-1111), or
// at the end.
//
// k_table_next, one workgroup per tile of 1024 blocks: the tile and the next
// three are staged in LDS as offsets from the tile's start, and for every block
// of the tile the table that would start there is found by a binary search:
// its end (tnext, as a distance; 0 when it lies past the staged span), its
// data bytes and its filter window count (FilterBlockBuilder emits one filter
// per 2 KiB of block start offsets up to the data end, plus one for the keys
// pending at Finish, table/filter_block.cc:21-49).  Per tile also its bytes
// and first forced block, and per block its offset in the tile, for tables
// longer than the span.
//
// k_cut_tables, one workgroup: the real tables are the chain of tnext from
// block 0.  Waves 1-15 stage tnext in LDS, 32768 blocks at a time, while
// wave 0 follows the chain through the chunk before (one LDS read a table);
// a table longer than the span is followed over the tiles' totals.  Then the
// tables' data ends and first filter windows (numbered across tables).
//
// k_table_layout, one workgroup per table: each block's table and offset in
// it, its index entry (size, then offset in the index block); each filter
// window's key count, first entry and filter size (then offset in the
// table's filter data); the table's size, which one scan over the tables
// turns into its offset in the output.

constexpr u32 kTnThreads = 1024;
constexpr u32 kTnAhead = 3;                             // tiles staged after the tile
constexpr u32 kTnSpan = (1 + kTnAhead) * kTableTile;  // blocks staged
constexpr u32 kTnPer = kTnSpan / kTnThreads;  // staged blocks per thread
constexpr u32 kTnStarts = kTableTile / kTnThreads;
constexpr u64 kTcNone = ~0ull;

__device__ __forceinline__ u64 table_windows(const EncParams& ep, u64 dend, u64 last) {
  return ep.bloom_bits ? max(dend >> 11, (last >> 11) + 1) : 0;
}

__global__ __launch_bounds__(kTnThreads) void k_table_next(const u32* nbo_dev, EncParams ep, EncBufs b, u32 size_cuts,
                                                             DevStatus* st) {
  __shared__ u64 sO[kTnSpan + 1];  // bytes of the staged blocks before each (with trailers), from the tile's start
  __shared__ u16 sNF[kTnSpan];     // the first forced block after each (kTnSpan: none staged)
  __shared__ u64 s_w[kTnThreads / 64];
  __shared__ u32 s_f[kTnThreads / 64];
  if (st->err) return;
  const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const u64 nbo = *nbo_dev;
  for (u64 T = blockIdx.x; T * kTableTile < nbo; T += gridDim.x) {
    const u64 c0 = T * kTableTile;
    const u32 send = (u32)(min(c0 + kTnSpan, nbo) - c0);  // staged blocks [0, send)
    const bool at_end = c0 + send == nbo;                   // y = send is then the stream's end
    const u32 x0 = tid * kTnPer;
    u64 v[kTnPer];
    u32 fmask = 0;
    u64 sum = 0;
#pragma unroll
    for (u32 k = 0; k < kTnPer; k++) {
      const u32 x = x0 + k;
      const bool real = x < send;
      v[k] = real ? (u64)b.ob_bytes[c0 + x] + 5 : 0ull;
      fmask |= (real && b.ob_forced[c0 + x] ? 1u : 0u) << k;
      sum += v[k];
    }
    const u64 incl = wave_incl_sum(sum);
    if (lane == 63) s_w[wave] = incl;
    // the first forced block of each thread, and after it (suffix minimum)
    u32 m = fmask ? x0 + (u32)__builtin_ctz(fmask) : kTnSpan;
#pragma unroll
    for (u32 d = 1; d < 64; d <<= 1) {
      const u32 o = (u32)__shfl_down((int)m, d, 64);
      if (lane + d < 64) m = min(m, o);
    }
    if (lane == 0) s_f[wave] = m;
    __syncthreads();
    u64 ex = incl - sum;
    u32 later = kTnSpan;
#pragma unroll
    for (u32 w = 0; w < kTnThreads / 64; w++) {
      if (w < wave) ex += s_w[w];
      if (w > wave) later = min(later, s_f[w]);
    }
    const u32 mnext = (u32)__shfl_down((int)m, 1, 64);
    u32 nxt = lane < 63 ? min(mnext, later) : later;
#pragma unroll
    for (u32 k = 0; k < kTnPer; k++) {
      sO[x0 + k] = ex;
      ex += v[k];
    }
    if (tid == kTnThreads - 1) sO[kTnSpan] = ex;
#pragma unroll
    for (int k = kTnPer - 1; k >= 0; k--) {
      sNF[x0 + k] = (u16)nxt;
      if ((fmask >> k) & 1) nxt = x0 + k;
    }
    __syncthreads();
    const u32 tn = min(send, kTableTile);  // table starts in this tile
    if (tid == 0) {
      b.tbytes[T] = sO[tn];
      const u32 f = (fmask & 1) ? 0u : (u32)sNF[0];
      b.tforced[T] = f < tn ? (u32)(c0 + f) : 0xffffffffu;
    }
    // P(y) = (y is the stream's end || O[y] >= target): monotone in y, on y
    // in (x, send] (sO[send] is the staged span's total)
#pragma unroll
    for (u32 k = 0; k < kTnStarts; k++) {
      const u32 x = tid + k * kTnThreads;
      if (x >= tn) break;
      const u64 ox = sO[x];
      b.torel[c0 + x] = ox;
      const u64 target = !size_cuts || ep.max_file_size > kTcNone - ox ? kTcNone : ox + ep.max_file_size;
      auto P = [&](u32 y) { return (at_end && y == send) || sO[y] >= target; };
      u32 e = 0xffffffffu;  // none in the span
      if (P(send)) {
        u32 lo = x + 1, hi = send;  // P(hi) holds; the first y with P(y) lies in [lo, hi]
        while (lo < hi) {
          const u32 mid = (lo + hi) >> 1;
          if (P(mid)) hi = mid;
          else lo = mid + 1;
        }
        e = lo;
      }
      const u32 fx = sNF[x];
      if (fx < send) e = min(e, fx);
      if (e != 0xffffffffu) {
        const u64 dend = sO[e] - ox, last = sO[e - 1] - ox;
        b.tnext[c0 + x] = (u16)(e - x);
        b.tdend[c0 + x] = dend;
        b.tnw[c0 + x] = (u32)table_windows(ep, dend, last);
      } else {
        b.tnext[c0 + x] = 0;  // past the span: followed over the tiles' totals
      }
    }
    __syncthreads();  // the next tile reuses the LDS
  }
}

// Exclusive scan over the workgroup of ld(i), i < n, one item per thread and
// round; out(i, prefix) for every item; returns the total.  Every item of a
// round is loaded before any of its outputs is written, so it may run in
// place.
template <u32 NT, typename T, typename Load, typename Out>
__device__ T wg_scan(u64 n, Load ld, Out out, T* s_w) {
  const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  T carry = 0;
  for (u64 base = 0; base < n; base += NT) {
    const u64 i = base + tid;
    const T v = i < n ? (T)ld(i) : T(0);
    const T incl = wave_incl_sum(v);
    if (lane == 63) s_w[wave] = incl;
    __syncthreads();
    T before = 0, tot = 0;
#pragma unroll
    for (u32 w = 0; w < NT / 64; w++) {
      const T x = s_w[w];
      if (w < wave) before += x;
      tot += x;
    }
    if (i < n) out(i, carry + before + incl - v);
    carry += tot;
    __syncthreads();
  }
  return carry;
}

constexpr u32 kCtThreads = 1024;
constexpr u32 kCtChunk = 32768;  // tnext entries staged per step

// A wave-uniform value, moved to scalar registers.
__device__ __forceinline__ u64 uni64(u64 v) {
  return (u64)(u32)__builtin_amdgcn_readfirstlane((int)(u32)v) |
         ((u64)(u32)__builtin_amdgcn_readfirstlane((int)(u32)(v >> 32)) << 32);
}

// The end of the table starting at block t when it lies past k_table_next's
// span (so past the three tiles after t's): whole tiles are passed over by their
// totals until one holds the size limit or a forced block, which is then
// searched 64 blocks at a time.  Writes the table's data bytes and windows
// at t, as k_table_next does for the others.  Wave 0 only, t uniform.
__device__ u64 table_end_far(const EncParams& ep, const EncBufs& b, u64 nbo, u32 size_cuts, u64 t) {
  const u32 lane = threadIdx.x & 63;
  const u64 need = size_cuts ? ep.max_file_size : kTcNone;
  const u64 T0 = t / kTableTile;
  u64 T = T0 + 1 + kTnAhead;
  u64 acc = b.tbytes[T0] - b.torel[t];  // O[T * tile] - O[t]
  for (u64 k = T0 + 1; k < T; k++) acc += b.tbytes[k];
  u64 e = nbo, ob = acc;  // the end and O[e] - O[t]
  for (;;) {
    const u64 c0 = T * kTableTile;
    if (c0 >= nbo) break;  // the stream ends at a tile boundary
    const u64 tb = b.tbytes[T];
    const u64 cn = min(c0 + kTableTile, nbo);  // blocks [c0, cn); cn is the next tile's first block or the end
    const u32 tf = b.tforced[T];
    u64 se = kTcNone;
    if (need != kTcNone && acc + tb >= need) {
      // the first j in [c0, cn] with acc + O_T[j] >= need (O_T[cn] = tb):
      // the last block of each run of 64, then the run of the first that holds
      u64 j = c0 + 64ull * lane + 63;
      bool p = j >= cn || acc + b.torel[j] >= need;
      u64 mk = __ballot(p);
      if (!mk) {
        se = cn;
      } else {
        const u64 r = c0 + 64ull * (u64)__builtin_ctzll(mk);
        j = r + lane;
        p = j >= cn || acc + b.torel[j] >= need;
        mk = __ballot(p);  // lane 63 holds
        se = r + (u64)__builtin_ctzll(mk);
      }
      se = uni64(se);
    }
    const u64 fe = tf != 0xffffffffu ? (u64)tf : kTcNone;
    if (se != kTcNone || fe != kTcNone) {
      e = min(se, fe);
      ob = acc + (e < cn ? b.torel[e] : tb);
      break;
    }
    acc += tb;
    ob = acc;
    T++;
  }
  e = uni64(e);
  ob = uni64(ob);
  const u64 last = ob - ((u64)b.ob_bytes[e - 1] + 5);
  if (lane == 0) {
    b.tdend[t] = ob;
    b.tnw[t] = (u32)table_windows(ep, ob, last);
  }
  return e;
}

__global__ __launch_bounds__(kCtThreads) void k_cut_tables(const u32* nbo_dev, EncParams ep, EncBufs b, u32 nf_cap,
                                                             u32 nw_cap, u32 size_cuts, DevStatus* st) {
  __shared__ u16 sN[2][kCtChunk];  // tnext of a chunk
  __shared__ u64 s_w[kCtThreads / 64];
  __shared__ u64 s_nf;
  const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const u64 nbo = *nbo_dev;
  if (st->err) {
    // the job is retried or rejected: no tables, so that the table offset
    // scan reads no sizes of an earlier job
    if (tid == 0) {
      st->n_files_out = 0;
      st->n_windows = 0;
      b.f_first[0] = (u32)nbo;
      b.f_win[0] = 0;
      b.f_off[0] = 0;
    }
    return;
  }
  const u64 nsteps = (nbo + kCtChunk - 1) / kCtChunk;
  u64 t = 0, nf = 0;  // wave 0: the chain's next table start, tables so far
  for (u64 s = 0; s <= nsteps; s++) {
    if (wave > 0 && s < nsteps) {
      // stage chunk s, 8 entries (16 bytes) per load; the arrays are
      // 16-byte aligned and chunks start at multiples of 8
      const u64 c0 = s * kCtChunk;
      u16* dst = sN[s & 1];
      for (u32 q = tid - 64; q < kCtChunk / 8; q += kCtThreads - 64) {
        const u64 i = c0 + 8ull * q;
        if (i >= nbo) break;
        if (i + 8 <= nbo) {
          *reinterpret_cast<uint4*>(dst + 8 * q) = *reinterpret_cast<const uint4*>(b.tnext + i);
        } else {
          for (u32 k = 0; i + k < nbo; k++) dst[8 * q + k] = b.tnext[i + k];
        }
      }
    }
    if (wave == 0 && s > 0) {
      const u64 c0 = (s - 1) * kCtChunk, c1 = min(c0 + kCtChunk, nbo);
      const u16* src = sN[(s - 1) & 1];
      while (t < c1) {
        const u32 d = src[t - c0];
        const u64 e = d ? t + d : table_end_far(ep, b, nbo, size_cuts, t);
        if (lane == 0 && nf < nf_cap) b.f_first[nf] = (u32)t;
        nf++;
        t = uni64(e);
      }
    }
    __syncthreads();
  }
  if (tid == 0) s_nf = nf;
  __syncthreads();
  nf = s_nf;
  if (nf > nf_cap) {  // reported for the host's retry; the tables are not laid out
    if (tid == 0) {
      st->need_files = (u32)min<u64>(nf, 0xffffffffull);
      set_err(st, ERR_OVERFLOW, SITE_CAPACITY);
      st->n_files_out = 0;
      st->n_windows = 0;
      b.f_first[0] = (u32)nbo;
      b.f_win[0] = 0;
      b.f_off[0] = 0;
    }
    return;
  }
  // the tables' data ends, and their windows numbered across tables
  u64 dsum = 0;
  const u64 nwt = wg_scan<kCtThreads, u64>(
      nf,
      [&](u64 k) -> u64 {
        const u32 t0 = b.f_first[k];
        const u64 dend = b.tdend[t0];
        dsum += dend;
        b.f_dend[k] = dend;
        return b.tnw[t0];
      },
      [&](u64 k, u64 ex) { b.f_win[k] = (u32)ex; }, s_w);
  dsum = wave_last(wave_incl_sum(dsum));
  if (lane == 63) s_w[wave] = dsum;
  __syncthreads();
  if (tid == 0) {
    u64 tot = 0;
    for (u32 w = 0; w < kCtThreads / 64; w++) tot += s_w[w];
    st->data_bytes = tot;
    u64 nfo = nf, nwo = nwt;
    if (nwt > nw_cap) {
      st->need_windows = (u32)min<u64>(nwt, 0xffffffffull);
      set_err(st, ERR_OVERFLOW, SITE_CAPACITY);
      nfo = 0;
      nwo = 0;
    }
    st->n_files_out = (u32)nfo;
    st->n_windows = (u32)nwo;
    b.f_first[nfo] = (u32)nbo;
    b.f_win[nfo] = (u32)nwo;
  }
}

constexpr u32 kTlThreads = 512;

__global__ __launch_bounds__(kTlThreads) void k_table_layout(Kept k, const u32* nf_dev, EncParams ep, EncBufs b,
                                                               ScanArgs sa, DevStatus* st) {
  __shared__ u64 s_w[kTlThreads / 64];
  __shared__ u32 s_ticket;
  if (st->err) return;  // nothing in this kernel sets it: every workgroup returns, or none
  const u64 nf = *nf_dev;
  const u64 f0 = lbscan::take_ticket(sa, &s_ticket);
  if (nf == 0 && f0 == 0 && threadIdx.x == 0) {
    b.f_off[0] = 0;
    st->out_bytes = 0;
  }
  for (u64 f = f0; f < nf; f += gridDim.x) {
    const u32 b0 = b.f_first[f], b1 = b.f_first[f + 1];
    const u64 dend = b.f_dend[f];
    // block offsets in the table: the bytes of the blocks before, with their trailers
    wg_scan<kTlThreads, u64>(
        b1 - b0, [&](u64 j) -> u64 { return (u64)b.ob_bytes[b0 + j] + 5; },
        [&](u64 j, u64 ex) {
          b.ob_off[b0 + j] = ex;
          b.ob_file[b0 + j] = (u32)f;
        },
        s_w);
    // index entries: the key after block i (a separator from the next
    // block's first key, or the last key's successor), the block handle
    const u64 isum = wg_scan<kTlThreads, u64>(
        b1 - b0,
        [&](u64 j) -> u64 {
          const u64 i = b0 + j;
          const u64 off = b.ob_off[i];
          const u64 last = (u64)b.ob_first[i] + b.ob_n[i] - 1;
          const u32 kl = i + 1 < b1 ? index_sep(k.K, last, b.ob_first[i + 1], nullptr) : index_succ(k.K, last, nullptr);
          const u32 hl = varint_len(off) + varint_len(b.ob_bytes[i]);
          return 1 + varint_len(kl) + varint_len(hl) + kl + hl;
        },
        [&](u64 j, u64 ex) { b.ob_isz[b0 + j] = ex; }, s_w);
    const u64 ib = isum + 4ull * (b1 - b0) + 4;
    u64 fb = 0;
    if (ep.bloom_bits) {
      // Block i lies in window (its offset >> 11), which never decreases
      // with i.  The first block of a window writes the window (its keys run
      // to the first block of a later window, or the table's end); the last
      // block of a window writes the empty windows up to the next block's
      // window, or to the table's window count.
      const u32 w0 = b.f_win[f];
      const u64 nwin = b.f_win[f + 1] - w0;
      for (u64 i = b0 + threadIdx.x; i < b1; i += kTlThreads) {
        const u64 sw = b.ob_off[i] >> 11;
        const bool first = i == b0 || (b.ob_off[i - 1] >> 11) != sw;
        const u64 wn = i + 1 < b1 ? b.ob_off[i + 1] >> 11 : nwin;
        if (first) {
          u64 q = i + 1;
          while (q < b1 && (b.ob_off[q] >> 11) == sw) q++;
          const u32 end = q < b1 ? b.ob_first[q] : b.ob_first[b1 - 1] + b.ob_n[b1 - 1];
          const u32 keys = end - b.ob_first[i];
          b.w_keys[w0 + sw] = keys;
          b.w_first[w0 + sw] = b.ob_first[i];
          b.w_file[w0 + sw] = (u32)f;
          b.w_fsz[w0 + sw] = filter_bytes(keys, ep.bloom_bits);
        }
        for (u64 e = sw + 1; e < wn; e++) {
          b.w_keys[w0 + e] = 0;
          b.w_first[w0 + e] = 0xffffffffu;
          b.w_file[w0 + e] = (u32)f;
          b.w_fsz[w0 + e] = 0;
        }
      }
      __syncthreads();
      const u64 fbytes = wg_scan<kTlThreads, u64>(
          nwin, [&](u64 j) -> u64 { return b.w_fsz[w0 + j]; }, [&](u64 j, u64 ex) { b.w_fsz[w0 + j] = ex; }, s_w);
      fb = fbytes + 4ull * nwin + 5;
      if (threadIdx.x == 0) {
        b.f_fbytes[f] = fbytes;
        b.f_fbase[f] = dend;
      }
    }
    if (threadIdx.x == 0) b.f_ibytes[f] = ib;
    // the table's offset in the output: the sizes of the tables before it,
    // by look-back over the tables (one a tile); the last one also gives
    // the output's size
    if (threadIdx.x < 64) {
      const u64 size = dend + (ep.bloom_bits ? fb + 5 : 0) + meta_block_size(ep, dend, fb) + 5 + ib + 5 + 48;
      const u64 ex = lbscan::publish<u64>(sa, f, size);
      if (threadIdx.x == 0) {
        b.f_off[f] = ex;
        if (f + 1 == nf) {
          b.f_off[nf] = ex + size;
          st->out_bytes = ex + size;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// writers



// The internal key user_key || fixed64(tag) of entry q as NWMAX big-endian
// words (bytes past the key are zero); the planes hold the user key's words.
template <u32 NWMAX>
__device__ __forceinline__ void ikey_words_of(const u64 (&uk)[NWMAX], u32 ulen, u64 tag, u64 (&ik)[NWMAX]) {
  const u64 tb = __builtin_bswap64(tag);
  const u32 wi = ulen >> 3, sft = (ulen & 7) * 8;
#pragma unroll
  for (u32 w = 0; w < NWMAX; w++) {
    u64 x = uk[w];
    if (w == wi) x |= tb >> sft;
    if (sft && w == wi + 1) x |= tb << (64 - sft);
    ik[w] = x;
  }
}
template <u32 NWMAX>
__device__ __forceinline__ void ikey_words(const Planes& K, u64 q, u32 ulen, u64 tag, u64 (&ik)[NWMAX]) {
  u64 uk[NWMAX];
#pragma unroll
  for (u32 w = 0; w < NWMAX; w++) uk[w] = w < K.W ? K.P[(u64)w * K.stride + q] : 0;
  ikey_words_of<NWMAX>(uk, ulen, tag, ik);
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

// ---------------------------------------------------------------------------
// The data block writer's LDS path.  gfx950 stores any width to any LDS byte
// address, so an entry's key delta, restart word and value go into the stage
// a word or 16 bytes at a time, always exactly (a store never touches a byte
// that belongs to another entry).

typedef u16 __attribute__((aligned(1))) u16_una;
struct __attribute__((aligned(1))) U4una {
  u32 x, y, z, w;
};
__device__ __forceinline__ void lds_st16(u8* p, u16 v) { *reinterpret_cast<u16_una*>(p) = v; }
__device__ __forceinline__ void lds_st32(u8* p, u32 v) { *reinterpret_cast<u32_una*>(p) = v; }
__device__ __forceinline__ void lds_st64(u8* p, u64 v) { *reinterpret_cast<u64_una*>(p) = v; }
__device__ __forceinline__ void lds_st128(u8* p, const uint4& v) {
  *reinterpret_cast<U4una*>(p) = U4una{v.x, v.y, v.z, v.w};
}

template <u32 N>
__device__ __forceinline__ u64 sel_word(const u64 (&w)[N], u32 i) {
  u64 x = 0;
#pragma unroll
  for (u32 t = 0; t < N; t++)
    if (t == i) x = w[t];
  return x;
}

// The n <= 8N bytes held by little-endian words d, stored at p.
template <u32 N>
__device__ __forceinline__ void lds_put_words(u8* p, const u64 (&d)[N], u32 n) {
#pragma unroll
  for (u32 w = 0; w < N; w++)
    if (8 * w + 8 <= n) lds_st64(p + 8 * w, d[w]);
  const u32 t0 = n & ~7u;
  u64 last = sel_word(d, t0 >> 3);
  u8* q = p + t0;
  if (n & 4) {
    lds_st32(q, (u32)last);
    last >>= 32;
    q += 4;
  }
  if (n & 2) {
    lds_st16(q, (u16)last);
    last >>= 16;
    q += 2;
  }
  if (n & 1) *q = (u8)last;
}

// Bytes [sh, sh + 8N) of the big-endian key words ik as little-endian words.
template <u32 N>
__device__ __forceinline__ void key_delta_words(const u64 (&ik)[N], u32 sh, u64 (&d)[N]) {
  const u32 s = 8 * (sh & 7);
#pragma unroll
  for (u32 j = 0; j < N; j++) {
    const u32 q = (sh >> 3) + j;
    const u64 hi = sel_word(ik, q), lo = sel_word(ik, q + 1);
    d[j] = __builtin_bswap64((hi << s) | ((lo >> 1) >> (63 - s)));
  }
}

// 16 bytes of a value source (always followed by at least 15 more readable
// bytes: a table's restart array, trailer and footer, or the padded value
// buffer of luda_build_tables).
__device__ __forceinline__ uint4 gload_16(u64 addr) {
  const GMEM u8* p = gbytes(addr);
  return uint4{gload_u32(p), gload_u32(p + 4), gload_u32(p + 8), gload_u32(p + 12)};
}

// The vl < 16 bytes of v stored at p.
__device__ __forceinline__ void lds_put_short(u8* p, const uint4& v, u32 vl) {
  u64 lo = (u64)v.x | ((u64)v.y << 32), hi = (u64)v.z | ((u64)v.w << 32);
  if (vl & 8) {
    lds_st64(p, lo);
    lo = hi;
    p += 8;
  }
  if (vl & 4) {
    lds_st32(p, (u32)lo);
    lo >>= 32;
    p += 4;
  }
  if (vl & 2) {
    lds_st16(p, (u16)lo);
    lo >>= 16;
    p += 2;
  }
  if (vl & 1) *p = (u8)lo;
}

// A value of at most 8 chunks of 16 bytes copied by its own lane in two
// halves: the first kHoist chunks loaded before the block's other work so
// that their latency overlaps it, then (the rest loaded) the stores into the
// stage.  A value's later chunks mostly lie in cache lines the next entry's
// first chunks fetched: entries are adjacent in their input block.  Chunks
// are read from src + 16 t, up to 15 bytes past the value, which every value
// source allows.
constexpr u32 kHoist = 4;  // more chunks' registers would spill
template <u32 KH>
struct ValueRegs {
  sv::Chunk c[KH];
};
__device__ __forceinline__ sv::Chunk chunk_at(u64 addr) {
  const uint4 v = gload_16(addr);
  return sv::Chunk{{v.x, v.y, v.z, v.w}};
}
template <u32 KH>
__device__ __forceinline__ void load_value_lane(ValueRegs<KH>& r, u64 src, u32 vl) {
  const u32 nch = (vl + 15) >> 4;
#pragma unroll
  for (u32 t = 0; t < KH; t++)
    if (t < nch) r.c[t] = chunk_at(src + 16 * t);
}
// The vl <= 128 bytes from src stored at dst as whole aligned dwords, junk
// at both ends (sv::store_value); chunks past the first KH (in r) are loaded
// KH at a time.
template <u32 KH>
__device__ __forceinline__ void store_value_lane(u8* dst, const ValueRegs<KH>& r, u64 src, u32 vl) {
  const u32 at = (u32)(uintptr_t)dst;
  u32* d32 = reinterpret_cast<u32*>(dst - (at & 3));
  sv::store_value<KH>(at, vl, r.c, [&](u32 t) { return chunk_at(src + 16 * t); }, [&](u32 k, u32 w) { d32[k] = w; });
}

// Which value slot each lane of a round takes: entry lanes mark where their
// slots start, and a max-scan over the lanes carries each mark forward.
struct ChunkTab {
  u32 mark[64];
  u32 cst[64];  // first slot of the entry lane
  u32 off[64];  // its value's offset in the block
  u32 vl[64];
  u64 src[64];
};

// The values of a group of up to 64 entries (values longer than 8 chunks
// somewhere in the group) spread over the wave by 16-byte slots of the
// stage: a slot inside a value is one unaligned 16-byte load and one aligned
// LDS store (unaligned LDS stores cost several aligned ones); the partial
// slots at a value's two ends are stored byte-exact.  Reads stay within 15
// bytes past a value, which every value source allows.
__device__ __forceinline__ void copy_values_spread(u8* stage, ChunkTab& ct, bool act, u32 voff, u64 vsrc, u32 vl) {
  const u32 lane = threadIdx.x & 63;
  const u32 g = (u32)(uintptr_t)stage & 15;  // stage byte x is 16-byte aligned when (g + x) % 16 == 0
  const u32 nsl = act && vl ? sv::spread_slots(g, voff, vl) : 0;
  const u32 cincl = wave_incl_sum(nsl);
  const u32 C = wave_last(cincl);
  const u32 cst = cincl - nsl;
  ct.cst[lane] = cst;
  ct.off[lane] = voff;
  ct.vl[lane] = vl;
  ct.src[lane] = vsrc;
  u32 carry = 0;
  for (u32 c0 = 0; c0 < C; c0 += 64) {
    ct.mark[lane] = 0;
    wave_sync();
    if (nsl && cst >= c0 && cst < c0 + 64) ct.mark[cst - c0] = lane + 1;
    wave_sync();
    u32 m = ct.mark[lane];
    m = wave_incl_max(m);
    const u32 e = m ? m - 1 : carry;
    carry = wave_last(e);
    const u32 c = c0 + lane;
    if (c < C) {
      const int eo = (int)ct.off[e];
      const sv::Slot sl = sv::spread_slot(g, eo, (int)ct.vl[e], c - ct.cst[e]);
      const uint4 v = gload_16(ct.src[e] + (u32)(sl.lo - eo));
      if (sl.full) *reinterpret_cast<uint4*>(stage + sl.x) = v;
      else lds_put_short(stage + sl.lo, v, (u32)(sl.hi - sl.lo));
    }
    wave_sync();
  }
}

// The bytes of output block i (entries, restart array, type byte) assembled
// in the wave's LDS stage at byte mis: BlockBuilder::Add/Finish
// (table/block_builder.cc:62-114) with the fork's (offset << 8) | count
// restart words.  One lane per entry: header, key delta from the key words
// in registers, restart word, and (values of up to 128 bytes) the value.
template <u32 NWMAX>
__device__ __forceinline__ void assemble_block_lds(const Kept& k, const EncParams& ep, const EncBufs& b, u64 q0,
                                                   u32 cnt, u64 bytes, u8* stage, ChunkTab& ct, DevStatus* st) {
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
    u32 vl = 0, voff = 0, ulen = 0, sh = 0;
    u64 vsrc = 0;
    if (act) {
      ulen = K.klen[q];
      sh = rs ? 0 : b.shared[q];
      vl = k.vlen[q];
      vsrc = k.vptr[q];
      const u32 il = ulen + 8;
      voff = (u32)off + varint_len(sh) + varint_len(il - sh) + varint_len(vl) + il - sh;
    }
    // values first, as in assemble_block_pre
    const u32 maxch = wave_max(act ? (vl + 15) >> 4 : 0);
    if (maxch <= 8) {
      if (act) {
        ValueRegs<(NWMAX <= 3 ? 2 : kHoist)> vr;  // registers are short with the loaded-ahead path beside
        load_value_lane(vr, vsrc, vl);
        store_value_lane(stage + voff, vr, vsrc, vl);
      }
    } else {
      copy_values_spread(stage, ct, act, voff, vsrc, vl);
    }
    wave_sync();
    if (act) {
      const u32 il = ulen + 8;
      u64 ik[NWMAX];
      ikey_words<NWMAX>(K, q, ulen, K.tag[q], ik);
      u8* p = stage + off;
      p = enc_varint(p, sh);
      p = enc_varint(p, il - sh);
      p = enc_varint(p, vl);
      u64 d[NWMAX];
      key_delta_words<NWMAX>(ik, sh, d);
      lds_put_words<NWMAX>(p, d, il - sh);
      if (rs) lds_st32(stage + dend + 4ull * (j / R), ((u32)off << 8) | min(R, cnt - j));
    }
    base += wave_last(incl);
  }
  wave_sync();
  if (lane == 0) {
    // the last value's junk may have reached the first restart word, which
    // an earlier round wrote: entry 0 starts the block
    lds_st32(stage + dend, min(R, cnt));
    lds_st32(stage + dend + 4ull * ngroups, ngroups);
    if (base != dend) set_err(st, ERR_CORRUPT, SITE_CAPACITY);
    stage[bytes] = 0;  // kNoCompression
  }
  wave_sync();
}

// One lane's entry of a block (lane j < cnt holds entry q0 + j), loaded a
// block ahead of its assembly so that the loads overlap the block before.
template <u32 NWMAX>
struct EntryRegs {
  u32 sz, sh, ulen, vl;
  u64 vptr, tag;
  u64 uk[NWMAX];  // user-key words (zero past the planes)
};
template <u32 NWMAX>
__device__ __forceinline__ EntryRegs<NWMAX> load_entry(const Kept& k, const EncBufs& b, u32 R, u64 q0, u32 cnt) {
  const u32 j = threadIdx.x & 63;
  const Planes& K = k.K;
  const bool act = j < cnt;
  const u64 q = act ? q0 + j : q0;  // idle lanes read entry q0 and drop it
  const bool rs = (j % R) == 0;
  EntryRegs<NWMAX> e;
  const u32 szr = b.szr[q], sznr = b.sznr[q], sh = b.shared[q];
  e.ulen = K.klen[q];
  e.vl = k.vlen[q];
  e.vptr = k.vptr[q];
  e.tag = K.tag[q];
#pragma unroll
  for (u32 w = 0; w < NWMAX; w++) e.uk[w] = w < K.W ? K.P[(u64)w * K.stride + q] : 0;
  e.sz = act ? (rs ? szr : sznr) : 0;
  e.sh = rs ? 0 : sh;
  if (!act) e.vl = 0;
  return e;
}

__device__ __forceinline__ u32 wave_max_dpp(u32 v) { return (u32)__builtin_amdgcn_readlane((int)wave_incl_max(v), 63); }

// assemble_block_lds for a block of at most 64 entries whose entries were
// loaded ahead; with vspread false the values are already in vr
// (load_value_lane), else the wave spreads their chunks itself.
template <u32 NWMAX>
__device__ __forceinline__ void assemble_block_pre(const EncParams& ep, const EntryRegs<NWMAX>& e, u32 cnt, u64 bytes,
                                                   u8* stage, ChunkTab& ct, DevStatus* st, const ValueRegs<kHoist>& vr,
                                                   bool vspread) {
  const u32 j = threadIdx.x & 63;
  const u32 R = ep.interval;
  const u32 ngroups = (cnt + R - 1) / R;
  const u64 dend = bytes - 4ull * ngroups - 4;  // where the restart array starts
  const bool act = j < cnt;
  const bool rs = (j % R) == 0;
  const u32 incl = wave_incl_sum(e.sz);
  const u32 off = incl - e.sz;
  const u32 il = e.ulen + 8;
  const u32 voff = off + varint_len(e.sh) + varint_len(il - e.sh) + varint_len(e.vl) + il - e.sh;
  // values first: store_value_lane leaves junk around each value that the
  // headers, keys and restart words written after it replace
  if (!vspread) {
    if (act) store_value_lane(stage + voff, vr, e.vptr, e.vl);
  } else {
    copy_values_spread(stage, ct, act, voff, e.vptr, e.vl);
  }
  wave_sync();
  if (act) {
    u64 ik[NWMAX];
    ikey_words_of<NWMAX>(e.uk, e.ulen, e.tag, ik);
    u8* p = stage + off;
    p = enc_varint(p, e.sh);
    p = enc_varint(p, il - e.sh);
    p = enc_varint(p, e.vl);
    u64 d[NWMAX];
    key_delta_words<NWMAX>(ik, e.sh, d);
    lds_put_words<NWMAX>(p, d, il - e.sh);
    if (rs) lds_st32(stage + dend + 4ull * (j / R), (off << 8) | min(R, cnt - j));
  }
  const u32 total = wave_last(incl);
  if (j == 0) {
    lds_st32(stage + dend + 4ull * ngroups, ngroups);
    if (total != dend) set_err(st, ERR_CORRUPT, SITE_CAPACITY);
    stage[bytes] = 0;  // kNoCompression
  }
  wave_sync();
}

typedef u32 u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ u32 sel4(const u32 (&w)[4], u32 i) {
  return i == 0 ? w[0] : (i == 1 ? w[1] : (i == 2 ? w[2] : w[3]));
}

// Bytes [lo, hi) of a 4-byte word (lo, hi may fall outside 0..4) as a mask.
__device__ __forceinline__ u32 byte_span_mask(int lo, int hi) {
  lo = lo < 0 ? 0 : (lo > 4 ? 4 : lo);
  hi = hi < 0 ? 0 : (hi > 4 ? 4 : hi);
  if (hi <= lo) return 0;
  const u32 a = lo == 4 ? 0u : (0xffffffffu << (8 * lo));
  const u32 z = hi == 4 ? 0xffffffffu : ~(0xffffffffu << (8 * hi));
  return a & z;
}

// One CRC-32C slicing-by-4 step: the raw register after the four bytes of w
// (little-endian) with register x before them folded in by the caller.
__device__ __forceinline__ u32 slice4(const u32 (*sl)[256], u32 x) {
  return sl[3][x & 0xff] ^ sl[2][(x >> 8) & 0xff] ^ sl[1][(x >> 16) & 0xff] ^ sl[0][x >> 24];
}

// One CRC-32C slicing-by-8 step: the raw register after 8 bytes, given
// x = register ^ bytes 0..3 and y = bytes 4..7 (little-endian words).
__device__ __forceinline__ u32 slice8(const u32 (*sl)[256], u32 x, u32 y) {
  return sl[7][x & 0xff] ^ sl[6][(x >> 8) & 0xff] ^ sl[5][(x >> 16) & 0xff] ^ sl[4][x >> 24] ^ sl[3][y & 0xff] ^
         sl[2][(y >> 8) & 0xff] ^ sl[1][(y >> 16) & 0xff] ^ sl[0][y >> 24];
}
// The raw register of 16 bytes from a zero register.
__device__ __forceinline__ u32 crc16(const u32 (*sl)[256], const u32 (&w)[4]) {
  return slice8(sl, slice8(sl, w[0], w[1]) ^ w[2], w[3]);
}

// The message of L bytes at stage byte mis (mis < 16; the stage is 16-byte
// aligned and dst is congruent to it mod 16) copied to dst, followed by its
// masked crc32c::Value.  The message is cut into the stage's 16-byte pieces:
// the F whole ones below the message end are right-aligned into rows of 64
// (1024 bytes), each lane folding one piece per row (every row's piece on its
// own, then the rows joined) and writing it to dst with one 16-byte store;
// the lanes are then joined (within rows of 16 lanes by DPP, the four rows by
// their first lanes), and the last e < 16 bytes appended.  The ~0 initial
// value is folded into the message's first four bytes and the stage bytes
// before mis are zeroed.  sl holds the slicing-by-8 tables, z1024 and zt[j]
// copies of g_crc.zeros[8] and g_crc.zeros[j + 2].  Needs F >= 2 and
// mis + L + 16 <= 1024 * kCrcRows.
constexpr u32 kCrcRows = (kStage + 32 + 1023) / 1024;
__device__ __forceinline__ void crc_copy_out(const uint4* stage4, u32 mis, u32 L, u8* dst, const u32 (*sl)[256],
                                             const u32 (*z1024)[256], const u32 (*zt)[4][256]) {
  const u32 lane = threadIdx.x & 63;
  const u8* stage = reinterpret_cast<const u8*>(stage4);
  const u32 E = mis + L, F = E >> 4, e = E & 15;
  GMEM u8* gd = (GMEM u8*)dst;
  GMEM u32x4* d4 = (GMEM u32x4*)(dst - mis);
  const u32 rows = (F + 63) >> 6, slot0 = 64 * rows - F;
  u32 c[kCrcRows];
#pragma unroll
  for (u32 r = 0; r < kCrcRows; r++) {
    c[r] = 0;
    if (r < rows) {
      const int f = (int)(64 * r + lane) - (int)slot0;
      uint4 v = {0, 0, 0, 0};
      if (f >= 0) {
        v = stage4[f];
        if (f > 0 || mis == 0) d4[f] = u32x4{v.x, v.y, v.z, v.w};
      }
      u32 w[4] = {v.x, v.y, v.z, v.w};
      if (r < 2) {  // the first two pieces are in rows 0 and 1
#pragma unroll
        for (u32 t = 0; t < 4; t++) {
          const int lo = 4 * (int)t + (f == 1 ? 16 : 0);
          if (f == 0) w[t] &= byte_span_mask((int)mis - lo, 4);
          if (f == 0 || f == 1) w[t] ^= byte_span_mask((int)mis - lo, (int)mis + 4 - lo);
        }
      }
      c[r] = crc16(sl, w);
    }
  }
  // the last e bytes, right-aligned in a zero-filled piece (all lanes alike)
  u32 clast = 0;
  if (e) {
    const uint4 t = stage4[F];
    u32 w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (u32 i = 0; i < 4; i++) w[i] &= byte_span_mask(0, (int)e - 4 * (int)i);
    const u32 z = 16 - e, a = z >> 2, bsh = z & 3;
    u32 o[4];
#pragma unroll
    for (u32 i = 0; i < 4; i++) {
      const int u = (int)i - (int)a;
      const u32 hi = u >= 0 ? sel4(w, (u32)u) : 0u;
      const u32 lo = u >= 1 ? sel4(w, (u32)(u - 1)) : 0u;
      o[i] = bsh ? __builtin_amdgcn_alignbyte(hi, lo, 4 - bsh) : hi;
    }
    clast = crc16(sl, o);
  }
  u32 s = c[0];
#pragma unroll
  for (u32 r = 1; r < kCrcRows; r++)
    if (r < rows) s = crc_shift(z1024, s) ^ c[r];
  // lane l's pieces end 16 (63 - l) bytes before the last whole piece does:
  // join pairs 1, 2, 4, 8 lanes apart within rows of 16 (DPP row_shl), ...
  {
    u32 o = dpp_u32<0x101, 0xf>(s);
    if ((lane & 1) == 0) s = crc_shift(zt[0], s) ^ o;
    o = dpp_u32<0x102, 0xf>(s);
    if ((lane & 3) == 0) s = crc_shift(zt[1], s) ^ o;
    o = dpp_u32<0x104, 0xf>(s);
    if ((lane & 7) == 0) s = crc_shift(zt[2], s) ^ o;
    o = dpp_u32<0x108, 0xf>(s);
    if ((lane & 15) == 0) s = crc_shift(zt[3], s) ^ o;
  }
  // ... then the four rows from their first lanes
  const u32 s0 = (u32)__builtin_amdgcn_readlane((int)s, 0), s1 = (u32)__builtin_amdgcn_readlane((int)s, 16);
  const u32 s2 = (u32)__builtin_amdgcn_readlane((int)s, 32), s3 = (u32)__builtin_amdgcn_readlane((int)s, 48);
  s = crc_shift(zt[5], crc_shift(zt[4], s0) ^ s1) ^ (crc_shift(zt[4], s2) ^ s3);
  if (e) {
    for (u32 i = 0; i < (e >> 2); i++) s = slice4(sl, s);
    for (u32 i = 0; i < (e & 3); i++) s = (s >> 8) ^ sl[0][s & 0xff];
    s ^= clast;
  }
  const u32 crc = crc_mask(s ^ 0xffffffffu);
  if (mis && lane < 16 - mis) gd[lane] = stage[mis + lane];
  // the last e message bytes and the 4 CRC bytes after them in one store
  if (lane < e + 4) gd[16 * F - mis + lane] = lane < e ? stage[16 * F + lane] : (u8)(crc >> (8 * (lane - e)));
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
// Each wavefront writes data blocks i, i + S, i + 2S, ... (S = all waves of
// the grid); one workgroup of 16 waves per CU holds every CRC table in LDS.
// Blocks that fit are assembled in the wave's LDS stage and written out with
// 16-byte stores while their CRC is folded; larger ones (a single big value)
// are assembled in place.  The next block's metadata and entries (blocks of
// at most 64 entries) are loaded while the current one is assembled and
// checksummed, the metadata of the one after it too.
constexpr u32 kWbThreads = 1024;
constexpr u32 kWbWaves = kWbThreads / 64;
template <u32 NWMAX>
__global__ __launch_bounds__(kWbThreads) void k_write_blocks(Kept k, const u32* nbo_dev, EncParams ep, EncBufs b,
                                                              u8* out, u64 out_cap, DevStatus* st) {
  constexpr bool kFast = NWMAX > 0;
  constexpr bool kPre = kFast && NWMAX <= 3;  // entries loaded a block ahead (registers allow keys <= 16 bytes)
  constexpr u32 NWP = kFast ? NWMAX : 1;
  union WaveTab {
    ChunkTab ct;  // the LDS path
    ValueTab vt;  // the byte-at-a-time path (NWMAX = 0)
  };
  // One LDS object with the CRC tables first: their offsets stay below
  // 64 KiB, so a lookup's table and row go into the ds_read's immediate
  // offset and only the byte index into a register.
  struct Lds {
    u32 sl[8][256];     // slicing tables (the fast path uses all eight, the others four)
    u32 zr[4][256];     // 1024 zero bytes (the fast path) or 256 (the others)
    u32 zt[6][4][256];  // 16 << j zero bytes: the joins of crc_copy_out
    uint4 stage_all[kWbWaves][(kStage + 32) / 16];
    WaveTab wtab[kWbWaves];
  };
  __shared__ Lds lds;
  auto& sl = lds.sl;
  auto& zr = lds.zr;
  auto& zt = lds.zt;
  auto& stage_all = lds.stage_all;
  auto& wtab = lds.wtab;
  if (st->out_bytes > out_cap || st->err) return;
  for (u32 i = threadIdx.x; i < 8 * 256; i += blockDim.x) sl[i >> 8][i & 255] = g_crc.slice[i >> 8][i & 255];
  for (u32 i = threadIdx.x; i < 4 * 256; i += blockDim.x) zr[i >> 8][i & 255] = g_crc.zeros[kFast ? 8 : 6][i >> 8][i & 255];
  if (kFast)
    for (u32 i = threadIdx.x; i < 6 * 4 * 256; i += blockDim.x)
      zt[i >> 10][(i >> 8) & 3][i & 255] = g_crc.zeros[2 + (i >> 10)][(i >> 8) & 3][i & 255];
  __syncthreads();
  const u64 nbo = *nbo_dev;
  const u32 wave = threadIdx.x >> 6;
  const u32 R = ep.interval;
  const u64 S = (u64)gridDim.x * kWbWaves;
  u64 i = (u64)blockIdx.x * kWbWaves + wave;
  auto fits = [&](const BlockMeta& m, u64 d) {
    return kFast && m.bytes + 5 + 16 <= kStage && (d & 15) + m.bytes + 1 >= 32;
  };
  BlockMeta m0 = {0, 0, 0, 0, 0}, m1 = {0, 0, 0, 0, 0};
  u64 d0 = 0;
  EntryRegs<NWP> e0, e1;
  bool p0 = false;
  if (i < nbo) {
    m0 = block_meta(b, i);
    d0 = b.f_off[m0.file] + m0.off;
    p0 = kPre && fits(m0, (u64)out + d0) && m0.n <= 64;
    if (p0) e0 = load_entry<NWP>(k, b, R, m0.first, m0.n);
  }
  if (i + S < nbo) m1 = block_meta(b, i + S);
  for (; i < nbo; i += S) {
    // This block's values (entries loaded ahead, values of at most 8 chunks)
    // are loaded first and waited for once its headers and keys are laid
    // out; the next blocks' metadata and entries are loaded after that,
    // while this block is checksummed (registers for both at once would
    // spill).
    ValueRegs<kHoist> vr;
    bool vspread = true;
    if (p0) {
      vspread = wave_max_dpp((e0.vl + 15) >> 4) > 8;  // vl is 0 on idle lanes
      if (!vspread) load_value_lane(vr, e0.vptr, e0.vl);
    }
    u8* dst = out + d0;
    const u64 bytes = m0.bytes;
    const u32 mis = (u32)((u64)dst & 15);
    const bool fit0 = fits(m0, (u64)dst);
    u8* stage = reinterpret_cast<u8*>(stage_all[wave]);
    if constexpr (kFast) {
      if (fit0) {
        if (p0) assemble_block_pre<NWMAX>(ep, e0, m0.n, bytes, stage + mis, wtab[wave].ct, st, vr, vspread);
        else assemble_block_lds<NWMAX>(k, ep, b, m0.first, m0.n, bytes, stage + mis, wtab[wave].ct, st);
      }
    }
    u64 d1 = 0;
    BlockMeta m2 = {0, 0, 0, 0, 0};
    if (i + S < nbo) d1 = b.f_off[m1.file] + m1.off;
    if (i + 2 * S < nbo) m2 = block_meta(b, i + 2 * S);
    const bool p1 = kPre && i + S < nbo && fits(m1, (u64)out + d1) && m1.n <= 64;
    if (p1) e1 = load_entry<NWP>(k, b, R, m1.first, m1.n);
    if (fit0) {
      crc_copy_out(stage_all[wave], mis, (u32)bytes + 1, dst, sl, zr, zt);
    } else {
      if (!kFast && bytes + 5 + 16 <= kStage) {
        assemble_block<NWMAX>(k, ep, b, m0.first, m0.n, bytes,
                              LdsBuf{stage, mis, &wtab[wave].vt}, sl, zr, st);
        wave_copy_out(dst, stage_all[wave], mis, bytes + 5);
      } else {
        // the row shift of this path is 256 bytes: zr holds it only when !kFast
        assemble_block<NWMAX>(k, ep, b, m0.first, m0.n, bytes, GlobalBuf{dst}, sl, kFast ? g_crc.zeros[6] : zr,
                              st);
      }
    }
    wave_sync();
    m0 = m1;
    d0 = d1;
    m1 = m2;
    e0 = e1;
    p0 = p1;
  }
}

// Where each kept entry's value lands in the output (for results kept
// decoded): one wavefront per output block walks its entries as
// assemble_block lays them out, entry header (three varints), key delta,
// value, and writes the value's device address.
__global__ __launch_bounds__(256) void k_value_addrs(Kept k, const u32* nbo_dev, EncParams ep, EncBufs b, u8* out,
                                                     u64* vaddr) {
  const u64 nbo = *nbo_dev;
  const u32 lane = threadIdx.x & 63;
  const u32 R = ep.interval;
  for (u64 i = (u64)blockIdx.x * 4 + (threadIdx.x >> 6); i < nbo; i += (u64)gridDim.x * 4) {
    const BlockMeta m = block_meta(b, i);
    const u64 base_addr = (u64)out + b.f_off[m.file] + m.off;
    u64 base = 0;
    for (u32 j0 = 0; j0 < m.n; j0 += 64) {
      const u32 j = j0 + lane;
      const bool act = j < m.n;
      const u64 q = (u64)m.first + j;
      const bool rs = (j % R) == 0;
      const u32 sz = act ? (rs ? b.szr[q] : b.sznr[q]) : 0;
      const u32 incl = wave_incl_scan(sz);
      if (act) {
        const u32 il = k.K.klen[q] + 8;
        const u32 sh = rs ? 0 : b.shared[q];
        const u32 vl = k.vlen[q];
        const u64 off = base + incl - sz;
        vaddr[q] = base_addr + off + varint_len(sh) + varint_len(il - sh) + varint_len(vl) + (il - sh);
      }
      base += wave_last(incl);
    }
  }
}

// util/hash.cc over a user key held as big-endian words uk (klen <= 8 KW).
template <u32 KW>
__device__ __forceinline__ u32 hash_words(const u64 (&uk)[KW], u32 klen, u32 seed) {
  const u32 m = 0xc6a4a793u;
  u32 h = seed ^ (klen * m);
  u32 i = 0;
#pragma unroll
  for (u32 w = 0; w < KW; w++) {
#pragma unroll
    for (u32 half = 0; half < 2; half++) {
      if (i + 4 <= klen) {
        h += __builtin_bswap32(half ? (u32)uk[w] : (u32)(uk[w] >> 32));
        h *= m;
        h ^= (h >> 16);
        i += 4;
      }
    }
  }
  const u32 rest = klen - i;
  if (rest) {
    const u64 word = sel_word(uk, i >> 3);
    const u32 half = (i & 4) ? (u32)word : (u32)(word >> 32);  // bytes i.. in its high bytes
    const u32 b0 = half >> 24, b1 = (half >> 16) & 0xff, b2 = (half >> 8) & 0xff;
    if (rest == 3) h += b2 << 16;
    if (rest >= 2) h += b1 << 8;
    h += b0;
    h *= m;
    h ^= (h >> 24);
  }
  return h;
}

// What a wave needs of filter window w before it can set the window's bits:
// its key count, first entry, and where its filter goes in the output.
struct WinHead {
  u32 n, first;
  u64 dst;
};

// One wavefront per filter window with keys: bloom bits with LDS atomics
// (BloomFilterPolicy::CreateFilter, util/bloom.cc:40-63).  A wave takes
// windows w, w + S, ...; the next window's head and user keys (its first 64)
// are loaded while the current window is built.  KW = 0 hashes the keys
// from the planes (any key width).
template <u32 KW>
__global__ __launch_bounds__(kThreads) void k_write_filters(Kept k, const u32* nw_dev, EncParams ep, EncBufs b,
                                                             u8* out, u64 out_cap, DevStatus* st) {
  __shared__ u32 bits_all[kWaves][kFBits / 4];
  if (st->out_bytes > out_cap || st->err) return;
  constexpr u32 KWP = KW ? KW : 1;
  const u64 nw = *nw_dev;
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const Planes& K = k.K;
  u32* lb = bits_all[wave];
  const u64 S = (u64)gridDim.x * kWaves;
  auto head = [&](u64 w) {
    WinHead h = {0, 0, 0};
    const u32 n = b.w_keys[w], first = b.w_first[w], f = b.w_file[w];
    const u64 fsz = b.w_fsz[w];
    if (n) {
      h.n = n;
      h.first = first;
      h.dst = b.f_off[f] + b.f_fbase[f] + fsz;
    }
    return h;
  };
  auto keys = [&](const WinHead& h, u64 (&uk)[KWP], u32& kl) {
    const u64 q = h.first + (lane < h.n ? lane : 0);
    kl = h.n ? K.klen[q] : 0;
#pragma unroll
    for (u32 w = 0; w < KWP; w++) uk[w] = (KW && h.n && w < K.W) ? K.P[(u64)w * K.stride + q] : 0;
  };
  u64 w = (u64)blockIdx.x * kWaves + wave;
  WinHead h0 = {0, 0, 0};
  u64 uk0[KWP];
  u32 kl0 = 0;
  if (w < nw) {
    h0 = head(w);
    keys(h0, uk0, kl0);
  }
  for (; w < nw; w += S) {
    WinHead h1 = {0, 0, 0};
    u64 uk1[KWP];
    u32 kl1 = 0;
    if (w + S < nw) {
      h1 = head(w + S);
      keys(h1, uk1, kl1);
    }
    const u32 n = h0.n;
    if (n) {
      u64 nbits = (u64)n * ep.bloom_bits;
      if (nbits < 64) nbits = 64;
      const u64 nbytes = (nbits + 7) / 8;
      nbits = nbytes * 8;
      if (nbytes > kFBits) {
        if (lane == 0) set_err(st, ERR_UNSUPPORTED, SITE_CAPACITY);
      } else {
        for (u32 i = lane; i < (nbytes + 3) / 4; i += 64) lb[i] = 0;
        wave_sync();
        for (u32 t = lane; t < n; t += 64) {
          u32 h;
          if (KW && t < 64) {
            h = hash_words<KWP>(uk0, kl0, 0xbc9f1d34u);
          } else {
            h = hash_ukey(K, h0.first + t, 0xbc9f1d34u);
          }
          const u32 delta = (h >> 17) | (h << 15);
          for (u32 j = 0; j < ep.bloom_k; j++) {
            const u32 bp = (u32)(h % nbits);
            atomicOr(&lb[bp >> 5], 1u << (bp & 31));
            h += delta;
          }
        }
        wave_sync();
        GMEM u8* dst = (GMEM u8*)(out + h0.dst);
        for (u32 i = lane; i < nbytes; i += 64) dst[i] = (u8)(lb[i >> 2] >> (8 * (i & 3)));
        if (lane == 0) dst[nbytes] = (u8)ep.bloom_k;
        wave_sync();
      }
    }
    h0 = h1;
#pragma unroll
    for (u32 x = 0; x < KWP; x++) uk0[x] = uk1[x];
    kl0 = kl1;
  }
}

__device__ __forceinline__ u8* put_fixed32(u8* p, u32 v) {
  st32(p, v);
  return p + 4;
}

// crc32c::Value of the L bytes at p with the rows split over the workgroup's
// NW waves: each wave folds a contiguous run of rows, and the runs are joined
// by appending to each the zero bytes of the rows after it.  Called by every
// thread of the workgroup; every thread gets the result.  zt holds the first
// nz zero-append maps.
template <u32 NW>
__device__ __forceinline__ u32 block_crc_split(u64 L, const GMEM u8* p, const u32 (*sl)[256],
                                               const u32 (*zt)[4][256], u32 nz, u32* part) {
  const u32 wave = threadIdx.x >> 6;
  const u64 rows = (L + 255) >> 8;
  const u64 per = (rows + NW - 1) / NW;
  const u64 r0 = min(rows, per * wave), r1 = min(rows, r0 + per);
  u32 s = wave_crc32c_rows_global(L, r0, r1, p, sl, zt[6], zt);
  s = crc_append_zeros_tab(s, 256 * (rows - r1), zt, nz);
  if ((threadIdx.x & 63) == 0) part[wave] = s;
  __syncthreads();
  u32 x = 0;
#pragma unroll
  for (u32 w = 0; w < NW; w++) x ^= part[w];
  __syncthreads();
  return x ^ 0xffffffffu;
}

constexpr u32 kTailThreads = 1024;
constexpr u32 kTailWaves = kTailThreads / 64;
constexpr u32 kTailZeros = 14;  // zero-append maps kept in LDS: up to 32 KiB of zeros

// Two workgroups per table when it has a filter, one otherwise: the first
// writes the filter block's offsets, trailer and CRC; the other the
// metaindex and index blocks with their trailers, the footer and the table's
// metadata.  The two blocks' CRCs (the long pieces of the tail) thus run on
// different CUs; everything either side needs is known from the layout.
template <u32 NWMAX>
__global__ __launch_bounds__(kTailThreads) void k_write_tail(Kept k, const u32* nf_dev, EncParams ep, EncBufs b, u8* out,
                                                          u64 out_cap, DevStatus* st) {
  __shared__ u32 sl[4][256];
  __shared__ u32 zl[kTailZeros][4][256];
  __shared__ u64 s_moff, s_msize;
  __shared__ u32 s_part[kTailWaves];
  if (st->out_bytes > out_cap || st->err) return;
  for (u32 i = threadIdx.x; i < 4 * 256; i += blockDim.x) sl[i >> 8][i & 255] = g_crc.slice[i >> 8][i & 255];
  for (u32 i = threadIdx.x; i < kTailZeros * 4 * 256; i += blockDim.x)
    zl[i >> 10][(i >> 8) & 3][i & 255] = g_crc.zeros[i >> 10][(i >> 8) & 3][i & 255];
  __syncthreads();
  const u64 nf = *nf_dev;
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const Planes& K = k.K;
  const u64 items = ep.bloom_bits ? 2 * nf : nf;
  for (u64 t = blockIdx.x; t < items; t += gridDim.x) {
    const bool filt = ep.bloom_bits && t < nf;  // this workgroup's part: the filter block
    const u64 f = ep.bloom_bits && !filt ? t - nf : t;
    const u32 b0 = b.f_first[f], b1 = b.f_first[f + 1];
    u8* base = out + b.f_off[f];
    const u64 dend = b.f_dend[f];
    u64 fbsize = 0;
    if (ep.bloom_bits) {
      const u32 w0 = b.f_win[f], w1 = b.f_win[f + 1];
      const u64 fbytes = b.f_fbytes[f];
      fbsize = fbytes + 4ull * (w1 - w0) + 5;
      if (filt) {
        // filter block: offsets array, array offset, base lg
        // (FilterBlockBuilder::Finish), type byte, CRC
        u8* fs = base + dend;
        for (u32 i = threadIdx.x; i < w1 - w0; i += blockDim.x)
          st32(fs + fbytes + 4ull * i, (u32)(b.w_fsz[w0 + i] - b.w_fsz[w0]));
        if (threadIdx.x == 0) {
          st32(fs + fbytes + 4ull * (w1 - w0), (u32)fbytes);
          fs[fbytes + 4ull * (w1 - w0) + 4] = 11;
          fs[fbsize] = 0;
        }
        __syncthreads();
        const u32 crc = block_crc_split<kTailWaves>(fbsize + 1, (const GMEM u8*)fs, sl, zl, kTailZeros, s_part);
        if (threadIdx.x == 0) st32(fs + fbsize + 1, crc_mask(crc));
        __syncthreads();
        continue;
      }
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
      const u64 isum = isize - 4ull * (b1 - b0) - 4;  // the entries' bytes, before the restart array
      for (u32 i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
        const u64 eoff = b.ob_isz[i];
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
    // type bytes, then the index block's CRC with its rows split over the
    // waves, the metaindex's by wave 0
    if (threadIdx.x == 0) {
      base[ioff + isize] = 0;
      base[s_moff + s_msize] = 0;
    }
    __syncthreads();
    {
      u8* is = base + ioff;
      const u32 crc = block_crc_split<kTailWaves>(isize + 1, (const GMEM u8*)is, sl, zl, kTailZeros, s_part);
      if (threadIdx.x == 0) st32(is + isize + 1, crc_mask(crc));
    }
    if (wave == 0) {
      u8* ms = base + s_moff;
      const u64 mrows = (s_msize + 1 + 255) >> 8;
      const u32 crc = wave_crc32c_rows_global(s_msize + 1, 0, mrows, (const GMEM u8*)ms, sl, zl[6], zl) ^ 0xffffffffu;
      if (lane == 0) st32(ms + s_msize + 1, crc_mask(crc));
    } else if (wave == 1) {
      // table metadata: Output::smallest / largest (db/db_impl.cc:1163-1166),
      // the keys' bytes one per lane
      const u64 qs = b.ob_first[b0];
      const u64 ql = (u64)b.ob_first[b1 - 1] + b.ob_n[b1 - 1] - 1;
      const u32 slen = K.klen[qs] + 8, llen = K.klen[ql] + 8;
      const u64 tags = K.tag[qs], tagl = K.tag[ql];
      u8* ks = b.keys + f * 2ull * ep.kw;
      u8* kl = ks + ep.kw;
      for (u32 t = lane; t < slen; t += 64) ks[t] = ikey_byte(K, qs, slen - 8, tags, t);
      for (u32 t = lane; t < llen; t += 64) kl[t] = ikey_byte(K, ql, llen - 8, tagl, t);
      if (lane == 0) {
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
        FileMeta fm;
        fm.offset = b.f_off[f];
        fm.size = (u64)(p - base);
        fm.nent = ql + 1 - qs;
        fm.slen = slen;
        fm.llen = llen;
        b.meta[f] = fm;
      }
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
  KLAUNCH(T, "bnext", k_bnext, std::min<u32>(grid_for(m_cap, kBnStarts), 8192), 256, 0, k, m_dev, ep, b, st);
}
void scan_blocks(KTimer& T, ScanState& S, Kept k, const u32* m_dev, u32 m_cap, EncBufs b, u32 nbo_cap,
                 DevStatus* st) {
  const ScanGuard g{nbo_cap, &st->need_blocks};
  lbscan::run<u32, u32>(T, "scan_blocks", S, lbscan::ArrLoad<u32>{b.bs}, (u32*)nullptr, m_dev, 0, m_cap,
                        &st->n_blocks_out, st, &g, BlockInfoEpi{k.forced, b, nbo_cap});
}
void launch_cut_tables(KTimer& T, const u32* nbo_dev, u32 nbo_cap, EncParams ep, EncBufs b, u32 nf_cap, u32 nw_cap,
                       bool size_cuts, DevStatus* st) {
  KLAUNCH(T, "table_next", k_table_next, grid_for(nbo_cap / kTableTile + 1, 1), kTnThreads, 0, nbo_dev, ep, b,
          size_cuts ? 1u : 0u, st);
  KLAUNCH(T, "cut_tables", k_cut_tables, 1, kCtThreads, 0, nbo_dev, ep, b, nf_cap, nw_cap, size_cuts ? 1u : 0u, st);
}
void launch_table_layout(KTimer& T, ScanState& S, Kept k, const u32* nf_dev, u32 nf_cap, EncParams ep, EncBufs b,
                         DevStatus* st) {
  if ((u64)nf_cap + 1 > S.tiles) throw_scan_capacity((u64)nf_cap + 1, S.tiles);
  ScanArgs sa;
  sa.words = S.words;
  sa.ticket = S.tickets ? S.ticket : nullptr;
  sa.epoch = scan_next_epoch(S, T.s);
  sa.cap = 0;
  sa.need = nullptr;
  sa.st = st;
  KLAUNCH(T, "table_layout", k_table_layout, grid_for(nf_cap, 1), kTlThreads, 0, k, nf_dev, ep, b, sa, st);
}
void launch_write_blocks(KTimer& T, Kept k, const u32* nbo_dev, u32 nbo_cap, EncParams ep, EncBufs b, u8* out,
                         u64 out_cap, DevStatus* st, u32 groups) {
  // one workgroup of 16 waves per CU (its LDS holds one); each wave loops
  const u32 g = std::min<u32>(grid_for(nbo_cap, kWbWaves), groups ? groups : 256);
  if (k.K.W <= 2)
    KLAUNCH(T, "write_blocks", k_write_blocks<3>, g, kWbThreads, 0, k, nbo_dev, ep, b, out, out_cap, st);
  else if (k.K.W <= 4)
    KLAUNCH(T, "write_blocks", k_write_blocks<5>, g, kWbThreads, 0, k, nbo_dev, ep, b, out, out_cap, st);
  else if (k.K.W <= 8)
    KLAUNCH(T, "write_blocks", k_write_blocks<9>, g, kWbThreads, 0, k, nbo_dev, ep, b, out, out_cap, st);
  else
    KLAUNCH(T, "write_blocks", k_write_blocks<0>, g, kWbThreads, 0, k, nbo_dev, ep, b, out, out_cap, st);
}
void launch_value_addrs(KTimer& T, Kept k, const u32* nbo_dev, u32 nbo_cap, EncParams ep, EncBufs b, u8* out,
                        u64* vaddr) {
  KLAUNCH(T, "value_addrs", k_value_addrs, std::min<u32>(grid_for(nbo_cap, 4), 8192), 256, 0, k, nbo_dev, ep, b, out,
          vaddr);
}
void launch_write_filters(KTimer& T, Kept k, const u32* nw_dev, u32 nw_cap, const u32* nf_dev, EncParams ep, EncBufs b,
                          u8* out, u64 out_cap, DevStatus* st) {
  (void)nf_dev;
  const u32 g = std::min<u32>(grid_for(nw_cap, kWaves), 2048);
  if (k.K.W <= 2)
    KLAUNCH(T, "write_filters", k_write_filters<2>, g, kThreads, 0, k, nw_dev, ep, b, out, out_cap, st);
  else if (k.K.W <= 4)
    KLAUNCH(T, "write_filters", k_write_filters<4>, g, kThreads, 0, k, nw_dev, ep, b, out, out_cap, st);
  else if (k.K.W <= 8)
    KLAUNCH(T, "write_filters", k_write_filters<8>, g, kThreads, 0, k, nw_dev, ep, b, out, out_cap, st);
  else
    KLAUNCH(T, "write_filters", k_write_filters<0>, g, kThreads, 0, k, nw_dev, ep, b, out, out_cap, st);
}
void launch_write_tail(KTimer& T, Kept k, const u32* nf_dev, u32 nf_cap, EncParams ep, EncBufs b, u8* out,
                       u64 out_cap, DevStatus* st) {
  const u32 g = grid_for(ep.bloom_bits ? 2ull * nf_cap : nf_cap, 1);
  if (k.K.W <= 2)
    KLAUNCH(T, "write_tail", k_write_tail<3>, g, kTailThreads, 0, k, nf_dev, ep, b, out, out_cap, st);
  else if (k.K.W <= 4)
    KLAUNCH(T, "write_tail", k_write_tail<5>, g, kTailThreads, 0, k, nf_dev, ep, b, out, out_cap, st);
  else if (k.K.W <= 8)
    KLAUNCH(T, "write_tail", k_write_tail<9>, g, kTailThreads, 0, k, nf_dev, ep, b, out, out_cap, st);
  else
    KLAUNCH(T, "write_tail", k_write_tail<0>, g, kTailThreads, 0, k, nf_dev, ep, b, out, out_cap, st);
}

// k_decode.hip — input side: footers and index blocks of the input tables,
// restart arrays of their data blocks, and the block decoder that turns every
// data block into key planes, tags and value references.
//
// Reference behaviour restated: Footer::DecodeFrom (table/format.cc:42-63),
// Table::Open (table/table.cc:38-82), ReadBlock's type handling
// (table/format.cc:65-154, CRCs not checked, as in the fork: :88), Block and
// DecodeEntry (table/block.cc:24-74, 224-251).  The fork's restart word is
// (offset << 8) | entries_in_group (table/block_builder.cc:62-74), which gives
// every restart group's entry count without parsing: that is what lets one
// lane decode one group with its output position known up front.
#include <algorithm>

#include "engine.h"
#include "scan.h"

namespace {

constexpr u32 kStage = 8192;  // bytes of a data block staged in LDS
typedef u32 u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ u32 wave_incl_scan_u32(u32 v) { return wave_incl_sum(v); }

// One input table: footer -> index block handle (Footer::DecodeFrom,
// table/format.cc:42-63; Table::Open, table/table.cc:38-82); returns its
// data block count (the index block's entries), 0 after an error.
__device__ u32 table_footer(const FileDesc& fd, FileInfo* info, DevStatus* st) {
  const u8* base = fd.base;
  const u64 size = fd.size;
  FileInfo fi = {0, 0, 0, 0};
  *info = fi;
  if (size < 48) {
    set_err(st, ERR_CORRUPT, SITE_FOOTER);
    return 0;
  }
  const u8* ft = base + size - 48;
  if (ld32(ft + 40) != 0x8b80fb57u || ld32(ft + 44) != 0xdb477524u) {
    set_err(st, ERR_CORRUPT, SITE_FOOTER);
    return 0;
  }
  u64 moff, msize, ioff, isize;
  const u8* p = ft;
  const u8* lim = ft + 48;
  if (!(p = dec_varint64(p, lim, &moff)) || !(p = dec_varint64(p, lim, &msize)) ||
      !(p = dec_varint64(p, lim, &ioff)) || !(p = dec_varint64(p, lim, &isize))) {
    set_err(st, ERR_CORRUPT, SITE_FOOTER);
    return 0;
  }
  if (ioff > size || isize > size || ioff + isize + 5 > size || isize < 4) {
    set_err(st, ERR_CORRUPT, SITE_INDEX);
    return 0;
  }
  if (base[ioff + isize] == 1) {  // kSnappyCompression
    set_err(st, ERR_UNSUPPORTED, SITE_COMPRESSED);
    return 0;
  }
  const u8* ib = base + ioff;
  const u32 nr = ld32(ib + isize - 4);
  if (nr > (isize - 4) / 4) {
    set_err(st, ERR_CORRUPT, SITE_INDEX);
    return 0;
  }
  u32 nb = nr;
  if (nr == 1 && (ld32(ib + isize - 8) & 0xff) == 0) nb = 0;  // empty table
  fi.ioff = ioff;
  fi.isize = isize;
  fi.nr = nr;
  *info = fi;
  return nb;
}

// The first launch of a read of tables, one workgroup: the job's status and
// the table descriptors from the pinned staging buffer (host memory, read in
// place), the footers (one thread per table), and each table's first data
// block: blk_base, the exclusive scan of the tables' block counts, with the
// total at blk_base[nfiles].
constexpr u32 kFilesThreads = 1024;
__global__ __launch_bounds__(kFilesThreads) void k_files_in(const FileDesc* src, const DevStatus* init, FileDesc* files,
                                                            u32 nfiles, FileInfo* info, u32* blk_base,
                                                            DevStatus* st) {
  __shared__ u32 s_w[kFilesThreads / 64];
  const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  {
    const u32* a = reinterpret_cast<const u32*>(init);
    u32* b = reinterpret_cast<u32*>(st);
    for (u32 i = tid; i < sizeof(DevStatus) / 4; i += kFilesThreads) b[i] = a[i];
  }
  __syncthreads();  // the status is set before any error is recorded in it
  u32 carry = 0;
  for (u32 f0 = 0; f0 < nfiles; f0 += kFilesThreads) {
    const u32 f = f0 + tid;
    u32 nb = 0;
    if (f < nfiles) {
      const FileDesc fd = src[f];
      files[f] = fd;
      nb = table_footer(fd, &info[f], st);
    }
    const u32 incl = wave_incl_sum(nb);
    if (lane == 63) s_w[wave] = incl;
    __syncthreads();
    u32 before = 0, tot = 0;
#pragma unroll
    for (u32 w = 0; w < kFilesThreads / 64; w++) {
      const u32 x = s_w[w];
      if (w < wave) before += x;
      tot += x;
    }
    if (f < nfiles) blk_base[f] = carry + before + incl - nb;
    carry += tot;
    __syncthreads();
  }
  if (tid == 0) blk_base[nfiles] = carry;
}

// 24 bytes of global memory as three little-endian words, and their bytes:
// the varints of an index entry or handle are decoded from registers after
// one round of loads instead of one load per byte.
struct Win24 {
  u64 w0, w1, w2;
};
__device__ __forceinline__ Win24 load_win24(const GMEM u8* p) { return {gload_u64(p), gload_u64(p + 8), gload_u64(p + 16)}; }
// (the word chosen by masks: a select between the fields can be turned into
// an indexed load from a stack copy of the window)
__device__ __forceinline__ u32 win_byte(const Win24& w, u32 i) {
  const u64 m0 = 0ull - (u64)(i < 8), m2 = 0ull - (u64)(i >= 16);
  const u64 x = (w.w0 & m0) | (w.w1 & ~(m0 | m2)) | (w.w2 & m2);
  return (u32)(x >> (8 * (i & 7))) & 0xff;
}
// GetVarint32Ptr / GetVarint64Ptr (util/coding.cc:121-165) on bytes [i, avail)
// of the window (avail <= 24); returns the position after it, or ~0u.
__device__ __forceinline__ u32 win_varint32(const Win24& w, u32 i, u32 avail, u32* v) {
  u32 r = 0;
  for (u32 shift = 0; shift <= 28 && i < avail; shift += 7) {
    const u32 b = win_byte(w, i++);
    r |= (b & 127) << shift;
    if (!(b & 128)) {
      *v = r;
      return i;
    }
  }
  return ~0u;
}
__device__ __forceinline__ u32 win_varint64(const Win24& w, u32 i, u32 avail, u64* v) {
  u64 r = 0;
  for (u32 shift = 0; shift <= 63 && i < avail; shift += 7) {
    const u64 b = win_byte(w, i++);
    r |= (b & 127) << shift;
    if (!(b & 128)) {
      *v = r;
      return i;
    }
  }
  return ~0u;
}

// The last f < n with a[f] <= j (a[0] = 0), and a[f] in *base, from LDS or
// from global memory: two functions whose loads all sit inside their own
// loops, so that no load can be merged into one through a generic pointer.
__device__ __forceinline__ u32 last_le_lds(const u32* a, u32 n, u32 j, u32* base) {
  u32 lo = 0, hi = n, v = 0;
  while (hi - lo > 1) {
    const u32 mid = (lo + hi) >> 1;
    const u32 x = a[mid];
    if (x <= j) {
      lo = mid;
      v = x;
    } else {
      hi = mid;
    }
  }
  *base = v;
  return lo;
}
__device__ __forceinline__ u32 last_le_global(const u32* a, u32 n, u32 j, u32* base) {
  const GMEM u32* g = (const GMEM u32*)a;
  u32 lo = 0, hi = n, v = 0;
  while (hi - lo > 1) {
    const u32 mid = (lo + hi) >> 1;
    const u32 x = g[mid];
    if (x <= j) {
      lo = mid;
      v = x;
    } else {
      hi = mid;
    }
  }
  *base = v;
  return lo;
}

// One thread per data block of any table: its index entry (a restart of its
// own, the index block's restart interval being 1, table/table_builder.cc:31)
// -> block handle, then the block's restart array -> entry count (the sum of
// the count bytes), with the structural checks the decoder relies on.  Also
// records the longest user key found at a restart point and the largest
// block and restart array, which size the decoder's LDS.  Thread j finds its
// table by a binary search of blk_base (staged in LDS when it is short).
// Reads that run past an entry stay inside its table: an index entry is
// followed by at least the restart array, the trailer and the footer (24
// bytes are read), and a data block by the index block and footer (16 bytes
// are read at a restart point).
constexpr u32 kIndexStage = 2048;

__global__ __launch_bounds__(256) void k_index(const FileDesc* files, const FileInfo* info, const u32* blk_base,
                                               u32 nfiles, u32 nb_cap, BlockDesc* blks, u32* blk_nent,
                                               u32* blk_ebase, u32* file_ebase, ScanArgs sa, DevStatus* st) {
  __shared__ u32 s_bb[kIndexStage];
  __shared__ u32 s_wsum[4], s_ex, s_ticket;
  const u32 total = blk_base[nfiles];
  if (blockIdx.x == 0 && threadIdx.x == 0) st->n_blocks_in = total > nb_cap ? 0 : total;
  if (total > nb_cap) {  // more handles than the tables' sizes allow
    if (blockIdx.x == 0 && threadIdx.x == 0) set_err(st, ERR_CORRUPT, SITE_INDEX);
    return;
  }
  const bool staged = nfiles + 1 <= kIndexStage;
  if (staged) {
    for (u32 i = threadIdx.x; i <= nfiles; i += blockDim.x) s_bb[i] = blk_base[i];
    __syncthreads();
  }
  u32 maxr = 0, maxn = 0, maxg = 0, maxe = 0;
  // Without blk_ebase, thread g takes blocks g, g + S, ...  With it (the
  // entry scan), workgroup w takes tiles w, w + G, ... of 256 blocks (by
  // ticket on worker slots) and looks back for the entries before each
  // (scan.h): a table's first block gives its first entry (file_ebase, the
  // empty tables before it the same), the last block the total (n_entries,
  // and file_ebase of the empty tables after it and of nfiles).
  const bool scan = blk_ebase != nullptr;
  const u64 ntile = ((u64)total + 255) / 256;
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const u64 step = scan ? (u64)gridDim.x * 256 : (u64)gridDim.x * blockDim.x;
  const u64 first = scan ? lbscan::take_ticket(sa, &s_ticket) * 256 + threadIdx.x
                         : (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u64 stop = scan ? (ntile ? ntile : 1) * 256 : (u64)total;
  auto table_of = [&](u32 j) {
    u32 fb;
    return staged ? last_le_lds(s_bb, nfiles, j, &fb) : last_le_global(blk_base, nfiles, j, &fb);
  };
  for (u64 jl = first; jl < stop; jl += step) {
    const u32 j = (u32)jl;
    u32 f = 0, c = 0;
    // one block; a failed check leaves it (`continue` ends the do-while)
    if (jl < total) do {
        blk_nent[j] = 0;
        u32 fb;
        if (staged) f = last_le_lds(s_bb, nfiles, j, &fb);
        else f = last_le_global(blk_base, nfiles, j, &fb);
        const u32 jf = j - fb;
        BlockDesc bd = {0, 0, f};
        blks[j] = bd;
        const u8* base = files[f].base;
        const u64 size = files[f].size;
        const FileInfo fi = info[f];
        const GMEM u8* ib = gbytes((u64)base + fi.ioff);
        const u64 iroff = fi.isize - 4 - 4ull * fi.nr;
        const u32 iw = gload_u32(ib + iroff + 4ull * jf);
        const u32 ioff = iw >> 8;
        if ((iw & 0xff) != 1 || ioff >= iroff) {
          set_err(st, ERR_CORRUPT, SITE_INDEX);
          continue;
        }
        const u64 iavail = iroff - ioff;  // bytes up to the restart array
        const Win24 e = load_win24(ib + ioff);
        const u32 ea = (u32)min<u64>(iavail, 24);
        u32 ish = 0, ins = 0, ivl = 0, x = 0;
        if ((x = win_varint32(e, 0, ea, &ish)) == ~0u || (x = win_varint32(e, x, ea, &ins)) == ~0u ||
            (x = win_varint32(e, x, ea, &ivl)) == ~0u || ish != 0 || iavail - x < (u64)ins + ivl) {
          set_err(st, ERR_CORRUPT, SITE_INDEX);
          continue;
        }
        const Win24 hw = load_win24(ib + ioff + x + ins);
        const u32 ha = min(ivl, 24u);
        u64 boff = 0, bsz = 0;
        u32 y = 0;
        if ((y = win_varint64(hw, 0, ha, &boff)) == ~0u || (y = win_varint64(hw, y, ha, &bsz)) == ~0u ||
            boff > size || bsz > size || boff + bsz + 5 > size || bsz > 0xffffffffull) {
          set_err(st, ERR_CORRUPT, SITE_INDEX);
          continue;
        }
        bd.addr = (u64)(base + boff);
        bd.size = (u32)bsz;
        blks[j] = bd;

        const GMEM u8* p = gbytes(bd.addr);
        const u32 n = (u32)bsz;
        if (n < 4) {
          set_err(st, p[n] == 1 ? ERR_UNSUPPORTED : ERR_CORRUPT, p[n] == 1 ? SITE_COMPRESSED : SITE_BLOCK_TRAILER);
          continue;
        }
        const u32 nr = gload_u32(p + n - 4);
        if (p[n] == 1) {
          set_err(st, ERR_UNSUPPORTED, SITE_COMPRESSED);
          continue;
        }
        if (nr > (n - 4) / 4) {
          set_err(st, ERR_CORRUPT, SITE_BLOCK_TRAILER);
          continue;
        }
        maxn = max(maxn, n);
        maxg = max(maxg, nr);
        if (nr == 0) continue;  // Block::NewIterator -> NewEmptyIterator
        const u32 roff = n - 4 - 4 * nr;
        u32 cnt_total = 0, prev = 0;
        bool bad = false;
        // restart words and the first bytes of their entries, two at a time
        auto one = [&](u32 kk, u32 w, u64 h0, u64 h1) {
          const u32 off = w >> 8, cnt = w & 0xff;
          if (roff == 0) {  // no entries at all
            if (cnt) bad = true;
            return;
          }
          if ((kk == 0 && off != 0) || (kk > 0 && off <= prev) || off >= roff || cnt == 0) bad = true;
          prev = off;
          cnt_total += cnt;
          // shared (0 at a restart) and key length: two varint32s, at most 10 bytes
          const Win24 hwin = {h0, h1, 0};
          const u32 avail = (u32)min<u64>(roff - off, 16);
          u32 sh = 0, ns = 0, z = 0;
          if (off < roff && (z = win_varint32(hwin, 0, avail, &sh)) != ~0u && win_varint32(hwin, z, avail, &ns) != ~0u &&
              sh == 0 && ns >= 8)
            maxr = max(maxr, ns - 8);
        };
        for (u32 k0 = 0; k0 < nr && !bad; k0 += 2) {
          const bool two = k0 + 1 < nr;
          const u32 wa = gload_u32(p + roff + 4 * k0);
          const u32 wb = two ? gload_u32(p + roff + 4 * (k0 + 1)) : 0;
          const bool ina = (wa >> 8) < roff, inb = two && (wb >> 8) < roff;
          const u64 a0 = ina ? gload_u64(p + (wa >> 8)) : 0, a1 = ina ? gload_u64(p + (wa >> 8) + 8) : 0;
          const u64 b0 = inb ? gload_u64(p + (wb >> 8)) : 0, b1 = inb ? gload_u64(p + (wb >> 8) + 8) : 0;
          one(k0, wa, a0, a1);
          if (two && !bad) one(k0 + 1, wb, b0, b1);
        }
        if (bad) {
          set_err(st, ERR_CORRUPT, SITE_BLOCK_TRAILER);
          continue;
        }
      blk_nent[j] = cnt_total;
      maxe = max(maxe, cnt_total);
      c = cnt_total;
    } while (0);
    if (!scan) continue;
    const u64 t = jl / 256;
    const u32 incl = wave_incl_sum(c);
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    u32 wpre = 0;
    for (u32 w = 0; w < wave; w++) wpre += s_wsum[w];
    const u32 agg = s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
    if (wave == 0) {
      const u32 ex = lbscan::publish<u32>(sa, t, agg);
      if (lane == 0) s_ex = ex;
    }
    __syncthreads();
    const u32 eb = s_ex + wpre + incl - c;
    if (jl < total) {
      blk_ebase[j] = eb;
      for (u32 g = j ? table_of(j - 1) + 1 : 0u; g <= f; g++) file_ebase[g] = eb;
    }
    if (jl + 1 == total || (total == 0 && jl == 0)) {
      const u32 n = s_ex + agg;
      for (u32 g = total ? table_of(total - 1) + 1 : 0u; g <= nfiles; g++) file_ebase[g] = n;
      st->n_entries = n;
    }
    __syncthreads();
  }
  if (maxr) atomicMax(&st->max_rulen, maxr);
  if (maxn) atomicMax(&st->max_blk, maxn);
  if (maxg) atomicMax(&st->max_nr, maxg);
  if (maxe) atomicMax(&st->max_ent, maxe);
}

// Byte sources for decode_block: the block staged in LDS, or read in place
// from global memory.  Each holds one kind of pointer only, so every read is
// a plain LDS (ds_*) or global load, never a generic (flat) one.
struct LdsBytes {
  const u8* p;
  __device__ __forceinline__ u32 at(u32 x) const { return p[x]; }
};
struct GlobalBytes {
  const u8* p;
  __device__ __forceinline__ u32 at(u32 x) const { return p[x]; }
};
template <typename Src>
__device__ __forceinline__ u32 src_ld32(const Src& s, u32 x) {
  return s.at(x) | (s.at(x + 1) << 8) | (s.at(x + 2) << 16) | (s.at(x + 3) << 24);
}
// GetVarint32Ptr (util/coding.cc:121-140) over positions [x, lim).
template <typename Src>
__device__ __forceinline__ bool src_varint32(const Src& s, u32& x, u32 lim, u32* v) {
  u32 r = 0;
  for (u32 shift = 0; shift <= 28 && x < lim; shift += 7) {
    const u32 c = s.at(x++);
    if (c & 128) {
      r |= (c & 127) << shift;
    } else {
      r |= c << shift;
      *v = r;
      return true;
    }
  }
  return false;
}

// Lane g decodes restart group g of one data block of n bytes whose first
// byte is at global address gp; the group's entries go to positions
// ebase + (entries of the groups before it).  kb is the lane's LDS key buffer.
template <typename Src>
__device__ __forceinline__ void decode_block(const Src& src, u32 n, u64 gp, u64 ebase, u8* kb, u32 KB,
                                             const Planes& K, u64* vptr, u32* vlen, DevStatus* st, u32& maxu) {
  const u32 lane = threadIdx.x & 63;
  const u32 nr = n >= 4 ? src_ld32(src, n - 4) : 0;
  if (nr == 0 || nr > (n - 4) / 4) return;
  const u32 roff = n - 4 - 4 * nr;
  for (u32 g0 = 0; g0 < nr; g0 += 64) {
    const u32 g = g0 + lane;
    const u32 w = g < nr ? src_ld32(src, roff + 4 * g) : 0;
    const u32 cnt = (g < nr && roff > 0) ? (w & 0xff) : 0;
    const u32 start = w >> 8;
    const u32 end = g + 1 < nr ? (src_ld32(src, roff + 4 * (g + 1)) >> 8) : roff;
    const u32 incl = wave_incl_scan_u32(cnt);
    const u32 chunk_total = wave_last(incl);
    u64 e = ebase + incl - cnt;
    if (cnt > 0) {
      u32 q = start;
      const u32 lim = end <= roff ? end : roff;
      u32 cur = 0;
      bool ok = start < end && end <= roff;
      for (u32 i = 0; ok && i < cnt; i++) {
        if (lim < q + 3) {  // DecodeEntry needs 3 bytes (table/block.cc:57)
          ok = false;
          break;
        }
        u32 sh = src.at(q), ns = src.at(q + 1), vl = src.at(q + 2);
        if ((sh | ns | vl) < 128) {
          q += 3;
        } else if (!src_varint32(src, q, lim, &sh) || !src_varint32(src, q, lim, &ns) ||
                   !src_varint32(src, q, lim, &vl)) {
          ok = false;
          break;
        }
        if (sh > cur || (i == 0 && sh != 0) || (u64)(lim - q) < (u64)ns + vl) {
          ok = false;
          break;
        }
        for (u32 t = 0; t < ns; t++) {
          const u32 pos = sh + t;
          if (pos < KB) kb[pos] = (u8)src.at(q + t);
        }
        cur = sh + ns;
        if (cur < 8) {
          set_err(st, ERR_UNSUPPORTED, SITE_SHORT_KEY);
          ok = false;
          break;
        }
        const u32 ulen = cur - 8;
        maxu = max(maxu, ulen);
        if (ulen <= 8 * K.W) {
          u64 tag = 0;
#pragma unroll
          for (u32 t = 0; t < 8; t++) tag |= (u64)kb[ulen + t] << (8 * t);
          for (u32 wd = 0; wd < K.W; wd++) {
            u64 word = 0;
#pragma unroll
            for (u32 t = 0; t < 8; t++) {
              const u32 pos = 8 * wd + t;
              word = (word << 8) | (pos < ulen ? kb[pos] : 0u);
            }
            K.P[wd * K.stride + e] = word;
          }
          K.klen[e] = ulen;
          K.tag[e] = tag;
        } else {
          // keep the entry in bounds (klen <= 8W); the host retries wider
          set_err(st, ERR_OVERFLOW, SITE_KEY_WIDTH);
          for (u32 wd = 0; wd < K.W; wd++) K.P[wd * K.stride + e] = 0;
          K.klen[e] = 8 * K.W;
          K.tag[e] = 0;
        }
        vptr[e] = gp + q + ns;
        vlen[e] = vl;
        q += ns + vl;
        e++;
      }
      if (!ok || q != lim) set_err(st, ERR_CORRUPT, SITE_GROUP_COUNT);
    }
    ebase += chunk_total;
  }
}

// ---------------------------------------------------------------------------
// Entry-parallel decoder.
//
// One wavefront per data block, one lane per entry.  Lane e finds its
// restart group from the groups' entry prefix (LDS), parses the headers from
// the group's restart point up to its own entry, and loads the bytes its
// entry stores (key bytes [shared, shared + non_shared)) into NW 64-bit
// big-endian words with a byte mask.  The full key of every entry then comes
// from one inclusive scan across the wave with the associative operator
//   (m1, v1) . (m2, v2) = (m1 | m2, (v1 & ~m2) | v2),
// i.e. every byte takes the value of the latest entry that stored it.  That
// is exactly BlockBuilder's prefix compression read back: byte j of key i is
// key i-1's byte j when j < shared_i, so it comes from the latest entry k <= i
// with shared_k <= j (table/block.cc:224-251).  Chunks of 64 entries carry
// the last lane's words into the next chunk.

// Byte windows over one data block: `win8(x)` is bytes [x, x+8) as a
// big-endian word (bytes past the block are don't-cares).
struct LdsWin {
  static constexpr bool kLds = true;
  const u8* stage;  // the LDS stage
  u32 mis;          // offset of the block's first byte in the stage
  // gfx950 reads 8 bytes at any LDS byte address in one access
  __device__ __forceinline__ u64 win8(u32 x) const {
    return __builtin_bswap64(*reinterpret_cast<const u64_una*>(stage + mis + x));
  }
  __device__ __forceinline__ u32 at(u32 x) const { return stage[mis + x]; }
  // bytes [x, x+4) little-endian (the stage is padded past the block)
  __device__ __forceinline__ u32 le32(u32 x) const { return *reinterpret_cast<const u32_una*>(stage + mis + x); }
};
struct GlobalWin {
  static constexpr bool kLds = false;
  const GMEM u8* p;
  u32 n;  // block bytes: nothing at or past n is read
  __device__ __forceinline__ u64 win8(u32 x) const {
    u64 v = 0;
    for (u32 t = 0; t < 8; t++) v = (v << 8) | (x + t < n ? p[x + t] : 0u);
    return v;
  }
  __device__ __forceinline__ u32 at(u32 x) const { return x < n ? p[x] : 0u; }
  __device__ __forceinline__ u32 le32(u32 x) const { return at(x) | at(x + 1) << 8 | at(x + 2) << 16 | at(x + 3) << 24; }
};
template <typename Src>
__device__ __forceinline__ u32 win_ld32(const Src& s, u32 x) {
  return __builtin_bswap32((u32)(s.win8(x) >> 32));
}
// DecodeEntry (table/block.cc:54-74) on [off, lim): false if malformed.
template <typename Src>
__device__ __forceinline__ bool win_entry_header(const Src& s, u32& off, u32 lim, u32& sh, u32& ns, u32& vl) {
  if (off > lim || lim - off < 3) return false;
  const u64 w = s.win8(off);
  const u32 b0 = (u32)(w >> 56), b1 = (u32)(w >> 48) & 0xff, b2 = (u32)(w >> 40) & 0xff;
  if (((b0 | b1 | b2) & 0x80) == 0) {
    sh = b0;
    ns = b1;
    vl = b2;
    off += 3;
  } else if (!src_varint32(s, off, lim, &sh) || !src_varint32(s, off, lim, &ns) ||
             !src_varint32(s, off, lim, &vl)) {
    return false;
  }
  return (u64)(lim - off) >= (u64)ns + vl;
}

// Scan state of decode_entries: the internal key's words and which of their
// bytes are known.  (0, 0) is the identity.
template <u32 NWMAX>
struct KeyBytes {
  u64 m[NWMAX], v[NWMAX];
  template <int CTRL, int ROW_MASK>
  __device__ __forceinline__ KeyBytes dpp() const {
    KeyBytes r;
#pragma unroll
    for (u32 w = 0; w < NWMAX; w++) {
      r.m[w] = dpp_u64<CTRL, ROW_MASK>(m[w]);
      r.v[w] = dpp_u64<CTRL, ROW_MASK>(v[w]);
    }
    return r;
  }
};

constexpr u32 kBadOff = 0xffff;  // eoff of an entry whose group failed its checks

// `mid` runs once between the restart-group walk and the entry loop (when
// nothing but the next block's loads is in flight).
template <u32 NWMAX, typename Src, typename Mid>
__device__ __forceinline__ void decode_entries(const Src& src, u32 n, u64 gp, u64 ebase, u32* gpre, u32 max_nr,
                                               u16* eoff, u32 ecap, const Planes& K, u64* vptr, u32* vlen,
                                               DevStatus* st, u32& maxu, Mid&& mid) {
  const u32 lane = threadIdx.x & 63;
  const u32 NW = K.W + 1;  // words of the internal key: user key words + the tag
  const u32 nr = n >= 4 ? win_ld32(src, n - 4) : 0;
  if (nr == 0 || nr > (n - 4) / 4) return;  // k_index reported it
  if (nr > max_nr) {  // not what k_index saw: gpre holds max_nr + 1 entries
    if (lane == 0) set_err(st, ERR_CORRUPT, SITE_BLOCK_TRAILER);
    return;
  }
  const u32 roff = n - 4 - 4 * nr;
  // entry prefix of the restart groups
  u32 total = 0;
  for (u32 g0 = 0; g0 < nr; g0 += 64) {
    const u32 g = g0 + lane;
    const u32 cnt = (g < nr && roff > 0) ? (win_ld32(src, roff + 4 * g) & 0xff) : 0;
    const u32 incl = wave_incl_scan_u32(cnt);
    if (g < nr) gpre[g] = total + incl - cnt;
    total += wave_last(incl);
  }
  if (lane == 0) gpre[nr] = total;
  __syncthreads();
  // A staged block whose entries fit eoff: one lane per restart group walks
  // the group once (DecodeEntry's checks, the first entry unshared, the last
  // one ending at the next restart point) and records every entry's offset;
  // the entry lanes then parse one header each instead of walking from their
  // group's restart point.
  const bool walked = Src::kLds && total <= ecap;
  if (walked) {
    for (u32 g0 = 0; g0 < nr; g0 += 64) {
      const u32 g = g0 + lane;
      if (g < nr) {
        const u32 start = win_ld32(src, roff + 4 * g) >> 8;
        const u32 end = g + 1 < nr ? (win_ld32(src, roff + 4 * (g + 1)) >> 8) : roff;
        const u32 lim = end <= roff ? end : roff;
        const u32 e0 = gpre[g], cnt = gpre[g + 1] - e0;
        bool ok = start < end && end <= roff;
        u32 off = start, i = 0;
        for (; ok && i < cnt; i++) {
          eoff[e0 + i] = (u16)off;
          // three one-byte varints: the next entry at off + 3 + ns + vl
          const u32 x = src.le32(off);
          if ((x & 0x808080u) == 0) {
            const u32 nx = off + 3 + ((x >> 8) & 0xff) + ((x >> 16) & 0xff);
            ok = nx <= lim && (i != 0 || (x & 0xff) == 0);
            off = nx;
          } else {
            u32 sh = 0, ns = 0, vl = 0;
            ok = win_entry_header(src, off, lim, sh, ns, vl);
            if (ok && i == 0 && sh != 0) ok = false;
            if (ok) off += ns + vl;
          }
        }
        if (ok && off != lim) ok = false;
        if (!ok)
          for (u32 t = i ? i - 1 : 0; t < cnt; t++) eoff[e0 + t] = (u16)kBadOff;
      }
    }
    __syncthreads();
  }
  mid();
  u64 cm[NWMAX], cv[NWMAX];  // words of the entry before the chunk
#pragma unroll
  for (u32 w = 0; w < NWMAX; w++) cm[w] = cv[w] = 0;
  u32 cilen = 0;
  bool bad = false;
  for (u32 e0 = 0; e0 < total; e0 += 64) {
    const u32 e = e0 + lane;
    const bool act = e < total;
    bool ok;
    u32 off, sh = 0, ns = 0, vl = 0;
    bool first = false;  // the first entry of its restart group
    if (walked) {
      off = act ? eoff[e] : kBadOff;
      ok = act && off != kBadOff;
      if (ok) ok = win_entry_header(src, off, n, sh, ns, vl);  // checked by the group's walk
    } else {
      u32 lo = 0, hi = nr;
      while (hi - lo > 1) {
        const u32 mid = (lo + hi) >> 1;
        if (gpre[mid] <= e) lo = mid;
        else hi = mid;
      }
      const u32 g = lo;
      const u32 r = e - gpre[g], cnt = gpre[g + 1] - gpre[g];
      first = r == 0;
      const u32 start = win_ld32(src, roff + 4 * g) >> 8;
      const u32 end = g + 1 < nr ? (win_ld32(src, roff + 4 * (g + 1)) >> 8) : roff;
      const u32 lim = end <= roff ? end : roff;
      ok = act && start < end && end <= roff;
      off = start;
      for (u32 i = 0; ok && i < r; i++) {
        ok = win_entry_header(src, off, lim, sh, ns, vl);
        off += ns + vl;
      }
      if (ok) ok = win_entry_header(src, off, lim, sh, ns, vl);
      if (ok && r == 0 && sh != 0) ok = false;
      if (ok && r + 1 == cnt && off + ns + vl != lim) ok = false;
    }
    const u32 koff = off, voff = off + ns;
    const u32 ilen = sh + ns;
    // this entry's bytes, positioned in the words of the internal key
    u64 m[NWMAX], v[NWMAX];
#pragma unroll
    for (u32 w = 0; w < NWMAX; w++) {
      m[w] = v[w] = 0;
      const u32 w0 = 8 * w;
      const u32 a = sh > w0 ? sh : w0;
      const u32 b = ilen < w0 + 8 ? ilen : w0 + 8;
      if (ok && w < NW && a < b) {
        const u32 lead = a - w0, tail = w0 + 8 - b;
        const u64 mask = (~0ull >> (8 * lead)) & (~0ull << (8 * tail));
        m[w] = mask;
        v[w] = (src.win8(koff + (a - sh)) >> (8 * lead)) & mask;
      }
    }
    // every byte from the latest entry that stored it
    {
      KeyBytes<NWMAX> x;
#pragma unroll
      for (u32 w = 0; w < NWMAX; w++) {
        x.m[w] = m[w];
        x.v[w] = v[w];
      }
      auto op = [](const KeyBytes<NWMAX>& a, const KeyBytes<NWMAX>& bb) {
        KeyBytes<NWMAX> r;
#pragma unroll
        for (u32 w = 0; w < NWMAX; w++) {
          r.v[w] = (a.v[w] & ~bb.m[w]) | bb.v[w];
          r.m[w] = a.m[w] | bb.m[w];
        }
        return r;
      };
      // with wide planes, a last chunk of few entries (blocks of large
      // values) scans only the lanes it uses; the carry out of a chunk is
      // read only by a next one.  (Narrow planes keep the one scan: the
      // extra code would take their registers past four waves per SIMD.)
      const u32 used = total - e0;
      if constexpr (NWMAX > 3) {
        if (used <= 8) x = wave_incl_scan_op_low<3>(x, op);
        else if (used <= 16) x = wave_incl_scan_op_low<4>(x, op);
        else if (used <= 32) x = wave_incl_scan_op_low<5>(x, op);
        else x = wave_incl_scan_op(x, op);
      } else {
        x = wave_incl_scan_op(x, op);
      }
#pragma unroll
      for (u32 w = 0; w < NWMAX; w++) {
        m[w] = x.m[w];
        v[w] = x.v[w];
      }
    }
#pragma unroll
    for (u32 w = 0; w < NWMAX; w++) {
      v[w] = (cv[w] & ~m[w]) | v[w];
      m[w] |= cm[w];
    }
    // shared may not exceed the previous key (DecodeEntry's key_.size() >= shared)
    const u32 up = dpp_u32<0x138, 0xf>(ilen);  // wave_shr:1: the previous lane's
    const u32 prev_ilen = lane == 0 ? cilen : up;
    if (ok && !first && sh > prev_ilen) ok = false;  // a group's first entry has sh = 0
    if (ok && ilen < 8) {
      set_err(st, ERR_UNSUPPORTED, SITE_SHORT_KEY);
      ok = false;
    }
    if (act) {
      const u64 q = ebase + e;
      if (ok && ilen <= 8 * NW) {
        const u32 ulen = ilen - 8;
        maxu = max(maxu, ulen);
        const u32 wi = ulen >> 3, sft = (ulen & 7) * 8;
        u64 x0 = 0, x1 = 0;
#pragma unroll
        for (u32 w = 0; w < NWMAX; w++) {
          if (w == wi) x0 = v[w];
          if (w == wi + 1) x1 = v[w];
        }
        const u64 be = sft ? (x0 << sft) | (x1 >> (64 - sft)) : x0;
#pragma unroll
        for (u32 w = 0; w + 1 < NWMAX; w++) {
          if (w < K.W) {
            const u32 nb = ulen > 8 * w ? min(ulen - 8 * w, 8u) : 0u;
            const u64 keep = nb ? (~0ull << (64 - 8 * nb)) : 0ull;
            K.P[w * K.stride + q] = v[w] & keep;
          }
        }
        K.klen[q] = ulen;
        K.tag[q] = __builtin_bswap64(be);
      } else {
        if (ok) {  // the host retries with planes as wide as this key
          set_err(st, ERR_OVERFLOW, SITE_KEY_WIDTH);
          maxu = max(maxu, ilen - 8);
        }
#pragma unroll
        for (u32 w = 0; w + 1 < NWMAX; w++)
          if (w < K.W) K.P[w * K.stride + q] = 0;
        K.klen[q] = 8 * K.W;
        K.tag[q] = 0;
      }
      vptr[q] = ok ? gp + voff : gp;
      vlen[q] = ok ? vl : 0;
    }
    bad |= act && !ok;
    // carry the chunk's last entry
#pragma unroll
    for (u32 w = 0; w < NWMAX; w++) {
      cm[w] = wave_last(m[w]);
      cv[w] = wave_last(v[w]);
    }
    cilen = wave_last(ilen);
  }
  if (bad) set_err(st, ERR_CORRUPT, SITE_GROUP_COUNT);
  __syncthreads();
}

// One wavefront per data block (see decode_entries).  The block is staged in
// LDS with 16-byte loads when it fits `stage` bytes, else read in place; the
// restart groups' entry prefix follows two stages.  The next block's first
// 8 KiB are loaded into registers while the current block's restart groups
// are walked, and put into the other stage before its entries are decoded,
// so a wave rarely waits for its bytes.  The descriptors of the wave's next
// 64 blocks are loaded together into LDS.
constexpr u32 kDescAhead = 64;
template <u32 NWMAX>
__global__ __launch_bounds__(64) void k_decode_entries(const BlockDesc* blks, u32 nb, const u32* blk_ebase, Planes K,
                                                       u64* vptr, u32* vlen, DevStatus* st, u32 stage, u32 max_nr,
                                                       u32 ecap, DecodeGuard guard) {
  if (guard.on) {
    // phase A's fields, final before this launch; none that the decoder
    // itself raises (max_ulen) is read
    if (st->err || st->site || st->n_entries > guard.n_cap || st->max_rulen > guard.max_rulen || st->max_nr > max_nr)
      return;
    nb = st->n_blocks_in;
  }
  extern __shared__ uint4 smem[];
  BlockDesc* dtab = reinterpret_cast<BlockDesc*>(smem + 2 * (stage / 16));
  u64* etab = reinterpret_cast<u64*>(dtab + kDescAhead);
  u32* gpre = reinterpret_cast<u32*>(etab + kDescAhead);
  u16* eoff = reinterpret_cast<u16*>(gpre + max_nr + 1);
  const u32 lane = threadIdx.x;
  constexpr u32 kBatch = 8;  // 16-byte pieces per lane in flight: a wave covers 8 KiB
  u32 maxu = 0;
  struct Span {
    u64 a0;     // the block's bytes lie in [a0, a0 + 16 n16) (16-byte pieces)
    u32 n16;    // 0 when not staged
    bool staged;
  };
  auto span_of = [&](const BlockDesc& d) {
    const u64 a0 = d.addr & ~15ull;
    const u64 a1 = (d.addr + d.size + 15) & ~15ull;
    const bool staged = a1 - a0 + 16 <= stage;
    return Span{a0, staged ? (u32)((a1 - a0) / 16) : 0u, staged};
  };
  auto stage_of = [&](u32 buf) { return reinterpret_cast<u64*>(smem + buf * (stage / 16)); };
  u64 lo[kBatch], hi[kBatch];
  // pieces [x0, x0 + 64 kBatch) of a staged block (idle lanes load piece 0)
  auto load_batch = [&](const Span& sp, u32 x0) {
    const GMEM u32x4* g16 = (const GMEM u32x4*)sp.a0;
#pragma unroll
    for (u32 t = 0; t < kBatch; t++) {
      const u32 x = x0 + 64 * t + lane;
      const u32x4 v = g16[x < sp.n16 ? x : 0];
      lo[t] = (u64)v[0] | ((u64)v[1] << 32);
      hi[t] = (u64)v[2] | ((u64)v[3] << 32);
    }
  };
  auto store_batch = [&](u64* s8, const Span& sp, u32 x0) {
#pragma unroll
    for (u32 t = 0; t < kBatch; t++) {
      const u32 x = x0 + 64 * t + lane;
      if (x < sp.n16) {
        s8[2 * x] = lo[t];
        s8[2 * x + 1] = hi[t];
      }
    }
  };
  const u64 G = gridDim.x;
  u64 b = blockIdx.x;
  // dtab[j], etab[j]: block b0 + j G (j < kDescAhead); the loop reloads them
  // when it reaches j = kDescAhead - 1, so the next block's is always there
  // (field by field: a struct copied between address spaces would go
  // through scratch)
  auto load_descs = [&](u64 b0) {
    const u64 bb = b0 + (u64)lane * G;
    u64 addr = 0, e = 0;
    u32 size = 0, file = 0;
    if (lane < kDescAhead && bb < nb) {
      addr = blks[bb].addr;
      size = blks[bb].size;
      file = blks[bb].file;
      e = blk_ebase[bb];
    }
    if (lane < kDescAhead) {
      dtab[lane].addr = addr;
      dtab[lane].size = size;
      dtab[lane].file = file;
      etab[lane] = e;
    }
    __syncthreads();
  };
  auto desc_at = [&](u32 jj) {
    BlockDesc d;
    d.addr = dtab[jj].addr;
    d.size = dtab[jj].size;
    d.file = dtab[jj].file;
    return d;
  };
  BlockDesc cur = {0, 0, 0};
  u64 ecur = 0;
  Span sc = {0, 0, false};
  bool held = false;  // lo/hi hold the first batch of `cur`, not yet in a stage
  u32 cb = 0;         // the stage of `cur`
  u32 j = 0;          // cur = dtab[j]
  if (b < nb) {
    load_descs(b);
    cur = desc_at(0);
    ecur = etab[0];
    sc = span_of(cur);
    if (sc.n16) {
      load_batch(sc, 0);
      held = true;
    }
  }
  for (; b < nb; b += G) {
    if (j == kDescAhead - 1) {
      load_descs(b);
      j = 0;
    }
    // the rest of the stage: a held first batch, the batches past 8 KiB
    u64* s8 = stage_of(cb);
    if (held) {
      store_batch(s8, sc, 0);
      held = false;
    }
    for (u32 x0 = 64 * kBatch; x0 < sc.n16; x0 += 64 * kBatch) {
      load_batch(sc, x0);
      store_batch(s8, sc, x0);
    }
    __syncthreads();
    const BlockDesc bd = cur;
    const u64 ebase = ecur;
    const bool staged = sc.staged;
    sc = {0, 0, false};
    j++;
    if (b + G < nb) {
      cur = desc_at(j);
      ecur = etab[j];
      sc = span_of(cur);
      if (sc.n16) {
        load_batch(sc, 0);
        held = true;
      }
    }
    u64* s8n = stage_of(cb ^ 1);
    auto mid = [&]() {
      if (held) {
        store_batch(s8n, sc, 0);
        held = false;
      }
    };
    const u8* gp = reinterpret_cast<const u8*>(bd.addr);
    const u32 n = bd.size;
    if (staged) {
      const LdsWin src{reinterpret_cast<const u8*>(s8), (u32)((u64)gp & 15)};
      decode_entries<NWMAX>(src, n, (u64)gp, ebase, gpre, max_nr, eoff, ecap, K, vptr, vlen, st, maxu, mid);
    } else {
      decode_entries<NWMAX>(GlobalWin{gbytes((u64)gp), n}, n, (u64)gp, ebase, gpre, max_nr, eoff, 0, K, vptr, vlen, st,
                            maxu, mid);
    }
    mid();  // when decode_entries returned early
    __syncthreads();
    cb ^= 1;
  }
  if (maxu) atomicMax(&st->max_ulen, maxu);
}

// Decoder for key planes wider than k_decode_entries handles (W > 8): one
// wavefront per data block, staged in LDS; lane g walks restart group g
// (entries known from the count byte, output position from a wave prefix sum
// of the counts), rebuilding each prefix-compressed key in its own LDS buffer.
// A block too large for the stage is read in place.
__global__ __launch_bounds__(64) void k_decode_groups(const FileDesc* files, const BlockDesc* blks, u32 nb,
                                               const u32* blk_ebase, Planes K, u64* vptr, u32* vlen,
                                               DevStatus* st) {
  extern __shared__ uint4 smem[];
  const u32 lane = threadIdx.x;
  const u32 KB = 8 * K.W + 8;
  u8* kb = reinterpret_cast<u8*>(smem) + kStage + 16 + lane * KB;
  u32 maxu = 0;
  for (u32 b = blockIdx.x; b < nb; b += gridDim.x) {
    const BlockDesc bd = blks[b];
    const u8* gp = reinterpret_cast<const u8*>(bd.addr);
    const u32 n = bd.size;
    const u64 ebase = blk_ebase[b];
    const u64 a0 = (u64)gp & ~15ull;
    const u64 a1 = ((u64)gp + n + 15) & ~15ull;
    if (a1 - a0 <= kStage + 16) {
      const uint4* g4 = reinterpret_cast<const uint4*>(a0);
      for (u32 x = lane; x < (a1 - a0) / 16; x += 64) smem[x] = g4[x];
      __syncthreads();
      const LdsBytes src{reinterpret_cast<const u8*>(smem) + ((u64)gp - a0)};
      decode_block(src, n, (u64)gp, ebase, kb, KB, K, vptr, vlen, st, maxu);
    } else {
      decode_block(GlobalBytes{gp}, n, (u64)gp, ebase, kb, KB, K, vptr, vlen, st, maxu);
    }
    __syncthreads();
  }
  if (maxu) atomicMax(&st->max_ulen, maxu);
}

// Decoded entries a previous call left resident (luda_job inputs produced
// with LUDA_JOB_KEEP_DECODED) copied into this job's arrays, in place of
// decoding their tables: entry e belongs to the source whose range holds it;
// plane words past the source's width are zero, as the decoder pads them.
__global__ void k_gather_decoded(const DecodedSrc* src, u32 nsrc, u64 n, Planes K, u64* vptr, u32* vlen) {
  GRID_STRIDE(e, n) {
    u32 lo = 0, hi = nsrc;
    while (hi - lo > 1) {
      const u32 mid = (lo + hi) >> 1;
      if (src[mid].dst <= e) lo = mid;
      else hi = mid;
    }
    const DecodedSrc& s = src[lo];
    const u64 j = s.first + (e - s.dst);
    for (u32 w = 0; w < K.W; w++) K.P[(u64)w * K.stride + e] = w < s.W ? s.P[(u64)w * s.stride + j] : 0;
    K.klen[e] = s.klen[j];
    K.tag[e] = s.tag[j];
    vptr[e] = s.vptr[j];
    vlen[e] = s.vlen[j];
  }
}

// Host entries -> kept stream (luda_build_tables).
__global__ void k_ingest(u64 n, const u8* keys, const u64* key_off, const u8* vals, const u64* val_off, Kept out,
                         DevStatus* st) {
  u32 maxu = 0;
  GRID_STRIDE(q, n) {
    const u64 k0 = key_off[q], k1 = key_off[q + 1];
    const u64 kl = k1 - k0;
    out.forced[q] = 0;
    out.vptr[q] = (u64)(vals + val_off[q]);
    out.vlen[q] = (u32)(val_off[q + 1] - val_off[q]);
    if (kl < 8) {
      set_err(st, ERR_UNSUPPORTED, SITE_SHORT_KEY);
      for (u32 wd = 0; wd < out.K.W; wd++) out.K.P[wd * out.K.stride + q] = 0;
      out.K.klen[q] = 0;
      out.K.tag[q] = 0;
      continue;
    }
    const u32 ulen = (u32)(kl - 8);
    maxu = max(maxu, ulen);
    if (ulen > 8 * out.K.W) {
      set_err(st, ERR_OVERFLOW, SITE_KEY_WIDTH);
      for (u32 wd = 0; wd < out.K.W; wd++) out.K.P[wd * out.K.stride + q] = 0;
      out.K.klen[q] = 8 * out.K.W;
      out.K.tag[q] = 0;
      continue;
    }
    const u8* k = keys + k0;
    for (u32 wd = 0; wd < out.K.W; wd++) {
      u64 word = 0;
      for (u32 t = 0; t < 8; t++) {
        const u32 pos = 8 * wd + t;
        word = (word << 8) | (pos < ulen ? k[pos] : 0u);
      }
      out.K.P[wd * out.K.stride + q] = word;
    }
    u64 tag = 0;
    for (u32 t = 0; t < 8; t++) tag |= (u64)k[ulen + t] << (8 * t);
    out.K.klen[q] = ulen;
    out.K.tag[q] = tag;
  }
  if (maxu) atomicMax(&st->max_ulen, maxu);
}

u32 grid_for(u64 n, u32 per_block) {
  u64 g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return (u32)g;
}

}  // namespace

void launch_files_in(KTimer& T, const FileDesc* src, const DevStatus* init, FileDesc* files, u32 nfiles,
                     FileInfo* info, u32* blk_base, DevStatus* st) {
  KLAUNCH(T, "files", k_files_in, 1, kFilesThreads, 0, src, init, files, nfiles, info, blk_base, st);
}
void launch_index(KTimer& T, const FileDesc* files, const FileInfo* info, const u32* blk_base, u32 nfiles,
                  u32 nb_cap, BlockDesc* blks, u32* blk_nent, DevStatus* st) {
  const ScanArgs none{nullptr, nullptr, 0, 0, nullptr, nullptr};
  KLAUNCH(T, "index", k_index, std::min<u32>(grid_for(nb_cap, 256), 1024), 256, 0, files, info, blk_base, nfiles, nb_cap, blks, blk_nent,
          (u32*)nullptr, (u32*)nullptr, none, st);
}
u64 index_scan_tiles(u32 nb_cap) { return ((u64)nb_cap + 255) / 256 + 1; }
void launch_index_entries(KTimer& T, ScanState& S, const FileDesc* files, const FileInfo* info, const u32* blk_base,
                          u32 nfiles, u32 nb_cap, BlockDesc* blks, u32* blk_nent, u32* blk_ebase, u32* file_ebase,
                          DevStatus* st) {
  if (index_scan_tiles(nb_cap) > S.tiles) throw_scan_capacity(index_scan_tiles(nb_cap), S.tiles);
  ScanArgs sa;
  sa.words = S.words;
  sa.ticket = S.tickets ? S.ticket : nullptr;
  sa.epoch = scan_next_epoch(S, T.s);
  sa.cap = 0;
  sa.need = nullptr;
  sa.st = st;
  KLAUNCH(T, "index", k_index, std::min<u32>(grid_for(nb_cap, 256), 1024), 256, 0, files, info, blk_base, nfiles,
          nb_cap, blks, blk_nent, blk_ebase, file_ebase, sa, st);
}
u64 decode_lds_bytes(u32 max_blk, u32 max_nr) {
  const u64 stage = (std::min<u64>(max_blk, 16384) + 15 + 16 + 15) & ~15ull;
  return 2 * stage + kDescAhead * (sizeof(BlockDesc) + 8) + 4ull * ((u64)max_nr + 1);
}
void launch_decode(KTimer& T, const FileDesc* files, const BlockDesc* blks, u32 nb, const u32* blk_ebase, Planes K,
                   u64* vptr, u32* vlen, DevStatus* st, u32 max_blk, u32 max_nr, u32 max_ent, u32 groups,
                   DecodeGuard guard) {
  if (!nb) return;
  if (guard.on && K.W > 8) launch_refused("decode (a guarded launch needs key planes of at most 8 words)", hipErrorInvalidValue);
  if (K.W <= 8) {
    // two stages of the largest block, its misalignment, and the 16 bytes an
    // 8-byte window may read past it; blocks over 16 KiB are read in place.
    // Then the descriptors of the next blocks, the restart groups' entry
    // prefix and, as far as 64 KiB allow, the entry offsets.
    const u32 stage = (u32)((std::min<u64>(max_blk, 16384) + 15 + 16 + 15) & ~15ull);
    size_t lds = decode_lds_bytes(max_blk, max_nr);
    if (lds > 65536) launch_refused("decode (restart arrays too long for LDS)", hipErrorInvalidValue);
    const u32 ecap = (u32)std::min<u64>((u64)max_ent, (65536 - lds) / 2);
    lds += ((2ull * ecap + 3) & ~3ull);
    // a few blocks per wavefront, so that the next descriptor's load overlaps
    const u32 grid = std::max<u32>(1, std::min<u32>((nb + 3) / 4, groups ? groups : 4096));
    if (K.W <= 2)
      KLAUNCH(T, "decode", k_decode_entries<3>, grid, 64, lds, blks, nb, blk_ebase, K, vptr, vlen, st, stage, max_nr,
              ecap, guard);
    else if (K.W <= 4)
      KLAUNCH(T, "decode", k_decode_entries<5>, grid, 64, lds, blks, nb, blk_ebase, K, vptr, vlen, st, stage, max_nr,
              ecap, guard);
    else
      KLAUNCH(T, "decode", k_decode_entries<9>, grid, 64, lds, blks, nb, blk_ebase, K, vptr, vlen, st, stage, max_nr,
              ecap, guard);
    return;
  }
  const u32 grid = nb < 65536 ? nb : 65536;
  // the stage and one key buffer per lane: 138 KiB for 2 KiB user keys, more
  // than the default limit of dynamic LDS per workgroup
  const size_t lds = kStage + 16 + 64 * (8 * K.W + 8);
  if (lds > 65536) {
    static bool raised = false;
    if (!raised) {
      const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_decode_groups),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      if (e != hipSuccess) launch_refused("decode (wide keys: LDS limit)", e);
      raised = true;
    }
  }
  KLAUNCH(T, "decode", k_decode_groups, grid, 64, lds, files, blks, nb, blk_ebase, K, vptr, vlen, st);
}
void launch_gather_decoded(KTimer& T, const DecodedSrc* src, u32 nsrc, u64 n, Planes K, u64* vptr, u32* vlen) {
  if (!n) return;
  KLAUNCH(T, "gather_decoded", k_gather_decoded, grid_for(n, 256), 256, 0, src, nsrc, n, K, vptr, vlen);
}
void launch_ingest(KTimer& T, u64 n, const u8* keys, const u64* key_off, const u8* vals, const u64* val_off, Kept out,
                   DevStatus* st) {
  if (!n) return;
  KLAUNCH(T, "ingest", k_ingest, grid_for(n, 256), 256, 0, n, keys, key_off, vals, val_off, out, st);
}

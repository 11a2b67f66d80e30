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
#include "engine.h"

namespace {

constexpr u32 kStage = 8192;  // bytes of a data block staged in LDS

__device__ __forceinline__ u32 wave_incl_scan_u32(u32 v) {
  const u32 lane = threadIdx.x & 63;
#pragma unroll
  for (u32 d = 1; d < 64; d <<= 1) {
    u32 o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// One thread per input table: footer -> index block handle.
__global__ void k_files(const FileDesc* files, u32 nfiles, FileInfo* info, u32* nblk, DevStatus* st) {
  GRID_STRIDE(f, nfiles) {
    nblk[f] = 0;
    const u8* base = files[f].base;
    const u64 size = files[f].size;
    FileInfo fi = {0, 0, 0, 0};
    info[f] = fi;
    if (size < 48) {
      set_err(st, ERR_CORRUPT, SITE_FOOTER);
      continue;
    }
    const u8* ft = base + size - 48;
    if (ld32(ft + 40) != 0x8b80fb57u || ld32(ft + 44) != 0xdb477524u) {
      set_err(st, ERR_CORRUPT, SITE_FOOTER);
      continue;
    }
    u64 moff, msize, ioff, isize;
    const u8* p = ft;
    const u8* lim = ft + 48;
    if (!(p = dec_varint64(p, lim, &moff)) || !(p = dec_varint64(p, lim, &msize)) ||
        !(p = dec_varint64(p, lim, &ioff)) || !(p = dec_varint64(p, lim, &isize))) {
      set_err(st, ERR_CORRUPT, SITE_FOOTER);
      continue;
    }
    if (ioff > size || isize > size || ioff + isize + 5 > size || isize < 4) {
      set_err(st, ERR_CORRUPT, SITE_INDEX);
      continue;
    }
    if (base[ioff + isize] == 1) {  // kSnappyCompression
      set_err(st, ERR_UNSUPPORTED, SITE_COMPRESSED);
      continue;
    }
    const u8* ib = base + ioff;
    const u32 nr = ld32(ib + isize - 4);
    if (nr > (isize - 4) / 4) {
      set_err(st, ERR_CORRUPT, SITE_INDEX);
      continue;
    }
    u32 nb = nr;
    if (nr == 1 && (ld32(ib + isize - 8) & 0xff) == 0) nb = 0;  // empty table
    fi.ioff = ioff;
    fi.isize = isize;
    fi.nr = nr;
    info[f] = fi;
    nblk[f] = nb;
  }
}

// One workgroup per table: each index entry (a restart of its own, the index
// block's restart interval being 1, table/table_builder.cc:31) -> block handle.
__global__ void k_index(const FileDesc* files, const FileInfo* info, const u32* blk_base, u32 nfiles,
                        BlockDesc* blks, DevStatus* st) {
  for (u32 f = blockIdx.x; f < nfiles; f += gridDim.x) {
    const u32 nb = blk_base[f + 1] - blk_base[f];
    if (nb == 0) continue;
    const u8* base = files[f].base;
    const u64 size = files[f].size;
    const FileInfo fi = info[f];
    const u8* ib = base + fi.ioff;
    const u64 roff = fi.isize - 4 - 4ull * fi.nr;
    for (u32 j = threadIdx.x; j < nb; j += blockDim.x) {
      BlockDesc bd = {0, 0, f};
      blks[blk_base[f] + j] = bd;
      const u32 w = ld32(ib + roff + 4ull * j);
      const u32 off = w >> 8;
      if ((w & 0xff) != 1 || off >= roff) {
        set_err(st, ERR_CORRUPT, SITE_INDEX);
        continue;
      }
      const u8* p = ib + off;
      const u8* lim = ib + roff;
      u32 sh, ns, vl;
      if (!(p = dec_varint32(p, lim, &sh)) || !(p = dec_varint32(p, lim, &ns)) || !(p = dec_varint32(p, lim, &vl)) ||
          sh != 0 || (u64)(lim - p) < (u64)ns + vl) {
        set_err(st, ERR_CORRUPT, SITE_INDEX);
        continue;
      }
      const u8* h = p + ns;
      const u8* hl = h + vl;
      u64 boff, bsz;
      if (!(h = dec_varint64(h, hl, &boff)) || !(h = dec_varint64(h, hl, &bsz)) || boff > size || bsz > size ||
          boff + bsz + 5 > size || bsz > 0xffffffffull) {
        set_err(st, ERR_CORRUPT, SITE_INDEX);
        continue;
      }
      bd.off = boff;
      bd.size = (u32)bsz;
      blks[blk_base[f] + j] = bd;
    }
  }
}

// One thread per data block: restart array -> entry count (sum of the count
// bytes), with the structural checks the decoder relies on.  Also records the
// longest user key found at a restart point.
__global__ void k_blkscan(const FileDesc* files, const BlockDesc* blks, const u32* nb_dev, u32* blk_nent,
                          DevStatus* st) {
  const u32 nb = *nb_dev;
  u32 maxr = 0;
  GRID_STRIDE(b, nb) {
    blk_nent[b] = 0;
    const BlockDesc bd = blks[b];
    const u8* p = files[bd.file].base + bd.off;
    const u32 n = bd.size;
    if (p[n] == 1) {
      set_err(st, ERR_UNSUPPORTED, SITE_COMPRESSED);
      continue;
    }
    if (n < 4) {
      set_err(st, ERR_CORRUPT, SITE_BLOCK_TRAILER);
      continue;
    }
    const u32 nr = ld32(p + n - 4);
    if (nr > (n - 4) / 4) {
      set_err(st, ERR_CORRUPT, SITE_BLOCK_TRAILER);
      continue;
    }
    if (nr == 0) continue;  // Block::NewIterator -> NewEmptyIterator
    const u32 roff = n - 4 - 4 * nr;
    u32 total = 0, prev = 0;
    bool bad = false;
    for (u32 k = 0; k < nr && !bad; k++) {
      const u32 w = ld32(p + roff + 4 * k);
      const u32 off = w >> 8, cnt = w & 0xff;
      if (roff == 0) {  // no entries at all
        if (cnt) bad = true;
        continue;
      }
      if ((k == 0 && off != 0) || (k > 0 && off <= prev) || off >= roff || cnt == 0) bad = true;
      prev = off;
      total += cnt;
      const u8* q = p + off;
      u32 sh, ns;
      const u8* r = dec_varint32(q, p + roff, &sh);
      if (r && (r = dec_varint32(r, p + roff, &ns)) && sh == 0 && ns >= 8) maxr = max(maxr, ns - 8);
    }
    if (bad) {
      set_err(st, ERR_CORRUPT, SITE_BLOCK_TRAILER);
      continue;
    }
    blk_nent[b] = total;
  }
  if (maxr) atomicMax(&st->max_rulen, maxr);
}

__global__ void k_file_ebase(const u32* blk_base, const u32* blk_ebase, u32 nfiles, const u32* n_entries,
                             u32* file_ebase) {
  GRID_STRIDE(f, nfiles + 1) {
    const u32 nb = blk_base[nfiles];
    const u32 b = blk_base[f];
    file_ebase[f] = b < nb ? blk_ebase[b] : *n_entries;
  }
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
    const u32 chunk_total = __shfl(incl, 63, 64);
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

// One wavefront per data block.  The block is staged in LDS with 16-byte
// loads; lane g then walks restart group g (entries known from the count
// byte, output position from a wave prefix sum of the counts), rebuilding
// each prefix-compressed key in its own LDS buffer and writing its planes.
// A block too large for the stage is read in place.
__global__ __launch_bounds__(64) void k_decode(const FileDesc* files, const BlockDesc* blks, u32 nb,
                                               const u32* blk_ebase, Planes K, u64* vptr, u32* vlen,
                                               DevStatus* st) {
  extern __shared__ uint4 smem[];
  const u32 lane = threadIdx.x;
  const u32 KB = 8 * K.W + 8;
  u8* kb = reinterpret_cast<u8*>(smem) + kStage + 16 + lane * KB;
  u32 maxu = 0;
  for (u32 b = blockIdx.x; b < nb; b += gridDim.x) {
    const BlockDesc bd = blks[b];
    const u8* gp = files[bd.file].base + bd.off;
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

void launch_files(KTimer& T, const FileDesc* files, u32 nfiles, FileInfo* info, u32* nblk, DevStatus* st) {
  KLAUNCH(T, "files", k_files, grid_for(nfiles, 64), 64, 0, files, nfiles, info, nblk, st);
}
void launch_index(KTimer& T, const FileDesc* files, const FileInfo* info, const u32* blk_base, u32 nfiles,
                  BlockDesc* blks, DevStatus* st) {
  KLAUNCH(T, "index", k_index, nfiles < 4096 ? nfiles : 4096, 256, 0, files, info, blk_base, nfiles, blks, st);
}
void launch_blkscan(KTimer& T, const FileDesc* files, const BlockDesc* blks, const u32* nb_dev, u32 nb_cap,
                    u32* blk_nent, DevStatus* st) {
  KLAUNCH(T, "blkscan", k_blkscan, grid_for(nb_cap, 256), 256, 0, files, blks, nb_dev, blk_nent, st);
}
void launch_file_ebase(KTimer& T, const u32* blk_base, const u32* blk_ebase, u32 nfiles, const u32* n_entries,
                       u32* file_ebase) {
  KLAUNCH(T, "file_ebase", k_file_ebase, grid_for(nfiles + 1, 64), 64, 0, blk_base, blk_ebase, nfiles, n_entries,
          file_ebase);
}
void launch_decode(KTimer& T, const FileDesc* files, const BlockDesc* blks, u32 nb, const u32* blk_ebase, Planes K,
                   u64* vptr, u32* vlen, DevStatus* st) {
  if (!nb) return;
  const size_t lds = kStage + 16 + 64 * (8 * K.W + 8);
  u32 grid = nb < 65536 ? nb : 65536;
  KLAUNCH(T, "decode", k_decode, grid, 64, lds, files, blks, nb, blk_ebase, K, vptr, vlen, st);
}
void launch_ingest(KTimer& T, u64 n, const u8* keys, const u64* key_off, const u8* vals, const u64* val_off, Kept out,
                   DevStatus* st) {
  if (!n) return;
  KLAUNCH(T, "ingest", k_ingest, grid_for(n, 256), 256, 0, n, keys, key_off, vals, val_off, out, st);
}

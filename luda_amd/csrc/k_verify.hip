// k_verify.hip — table verification: the masked CRC-32C stored after every
// block (the block's bytes and its type byte) is recomputed and compared, as
// ReadBlock does with ReadOptions::verify_checksums (table/format.cc:83-96;
// the fork never checks, :88), for every data block, the index block, the
// metaindex block and every block the metaindex names (the filter).
#include "engine.h"

namespace {

constexpr u32 kVThreads = 256;
constexpr u32 kVWaves = kVThreads / 64;
constexpr u32 kVZeros = 7;  // zero-append maps 0..6: the lane joins and the 256-byte row shift

// CRC tables into LDS; every thread of the workgroup calls it.
__device__ __forceinline__ void load_tables(u32 (*sl)[256], u32 (*zl)[4][256]) {
  for (u32 i = threadIdx.x; i < 4 * 256; i += blockDim.x) sl[i >> 8][i & 255] = g_crc.slice[i >> 8][i & 255];
  for (u32 i = threadIdx.x; i < kVZeros * 4 * 256; i += blockDim.x)
    zl[i >> 10][(i >> 8) & 3][i & 255] = g_crc.zeros[i >> 10][(i >> 8) & 3][i & 255];
  __syncthreads();
}

// Whether the masked CRC stored after block [p, p + n) and its type byte
// matches (the whole wave computes it; every lane gets the answer).
__device__ __forceinline__ bool wave_block_ok(u64 addr, u64 n, const u32 (*sl)[256], const u32 (*zl)[4][256]) {
  const GMEM u8* p = gbytes(addr);
  const u64 L = n + 1;
  const u32 s = wave_crc32c_rows_global(L, 0, (L + 255) >> 8, p, sl, zl[6], zl);
  return crc_mask(s ^ 0xffffffffu) == gload_u32(p + L);
}

__device__ __forceinline__ void report(DevStatus* st, u32 file) {
  if ((threadIdx.x & 63) == 0) {
    set_err(st, ERR_CORRUPT, SITE_CHECKSUM);
    atomicMin(&st->bad_file, file);
  }
}

// One wavefront per data block.  k_index leaves a block's address 0 when its
// handle failed a bounds check; such a block is skipped (already reported).
__global__ __launch_bounds__(kVThreads) void k_verify_blocks(const BlockDesc* blks, DevStatus* st) {
  __shared__ u32 sl[4][256];
  __shared__ u32 zl[kVZeros][4][256];
  load_tables(sl, zl);
  const u32 nb = st->n_blocks_in;
  const u64 waves = (u64)gridDim.x * kVWaves;
  for (u64 j = (u64)blockIdx.x * kVWaves + (threadIdx.x >> 6); j < nb; j += waves) {
    const u64 addr = blks[j].addr;
    const u32 size = blks[j].size;
    const u32 file = blks[j].file;
    if (addr == 0) continue;
    if (!wave_block_ok(addr, size, sl, zl)) report(st, file);
  }
}

// A BlockHandle (two varint64s, table/format.cc:23-33) whose block and
// trailer lie inside the table.
__device__ __forceinline__ const u8* handle_in(const u8* p, const u8* lim, u64 size, u64* off, u64* n) {
  if (!(p = dec_varint64(p, lim, off)) || !(p = dec_varint64(p, lim, n))) return nullptr;
  if (*off > size || *n > size || *off + *n + 5 > size) return nullptr;
  return p;
}

// One wavefront per table: footer, index block, metaindex block, and every
// block the metaindex names.  All lanes parse alike; only the CRCs are split.
__global__ __launch_bounds__(kVThreads) void k_verify_meta(const FileDesc* files, u32 nfiles, DevStatus* st) {
  __shared__ u32 sl[4][256];
  __shared__ u32 zl[kVZeros][4][256];
  load_tables(sl, zl);
  const u64 waves = (u64)gridDim.x * kVWaves;
  for (u64 f = (u64)blockIdx.x * kVWaves + (threadIdx.x >> 6); f < nfiles; f += waves) {
    const u8* base = files[f].base;
    const u64 size = files[f].size;
    bool ok = size >= 48;
    u64 moff = 0, msize = 0, ioff = 0, isize = 0;
    if (ok) {
      const u8* ft = base + size - 48;
      ok = ld32(ft + 40) == 0x8b80fb57u && ld32(ft + 44) == 0xdb477524u;
      const u8* p = ft;
      if (ok) ok = (p = handle_in(p, ft + 48, size, &moff, &msize)) != nullptr;
      if (ok) ok = handle_in(p, ft + 48, size, &ioff, &isize) != nullptr;
    }
    if (ok) ok = wave_block_ok((u64)(base + ioff), isize, sl, zl);
    if (ok) ok = msize >= 4 && wave_block_ok((u64)(base + moff), msize, sl, zl);
    if (ok) {
      // metaindex entries (Block::Iter over a block of restart interval 1)
      const u8* mb = base + moff;
      const u32 nr = ld32(mb + msize - 4);
      ok = nr <= (msize - 4) / 4;
      const u8* p = mb;
      const u8* lim = ok ? mb + msize - 4 - 4ull * nr : mb;
      for (u32 e = 0; ok && p < lim && e < 64; e++) {
        u64 sh, ns, vl, boff, bn;
        if (!(p = dec_varint64(p, lim, &sh)) || !(p = dec_varint64(p, lim, &ns)) ||
            !(p = dec_varint64(p, lim, &vl)) || (u64)(lim - p) < ns + vl) {
          ok = false;
          break;
        }
        const u8* vend = p + ns + vl;
        if (!handle_in(p + ns, vend, size, &boff, &bn)) {
          ok = false;
          break;
        }
        ok = wave_block_ok((u64)(base + boff), bn, sl, zl);
        p = vend;
      }
    }
    if (!ok) report(st, (u32)f);
  }
}

}  // namespace

void launch_verify(KTimer& T, const FileDesc* files, u32 nfiles, const BlockDesc* blks, u32 nb_cap, DevStatus* st) {
  const u32 gb = (u32)std::min<u64>(std::max<u64>((nb_cap + kVWaves - 1) / kVWaves, 1), 2048);
  KLAUNCH(T, "verify_blocks", k_verify_blocks, gb, kVThreads, 0, blks, st);
  const u32 gm = (u32)std::min<u64>(std::max<u64>((nfiles + kVWaves - 1) / kVWaves, 1), 2048);
  KLAUNCH(T, "verify_meta", k_verify_meta, gm, kVThreads, 0, files, nfiles, st);
}

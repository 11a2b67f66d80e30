// k_read.hip — the read side: decoded entries of whole tables written out as
// plain (internal key, value) arrays, in table order, as Table::Open and a
// full iteration (a TwoLevelIterator over Block::Iter, table/table.cc:38-214,
// table/block.cc:224-251) yield them.
#include "engine.h"

namespace {

// Output size of entry e's internal key (user key + 8-byte tag); item n of
// both size arrays is 0, so that exclusive scans over n + 1 items end with
// the totals.
__global__ void k_entry_out_sizes(Planes K, u64 n, u32* ksz, u32* vlen) {
  GRID_STRIDE(e, n + 1) {
    if (e < n) {
      ksz[e] = K.klen[e] + 8;
    } else {
      ksz[e] = 0;
      vlen[e] = 0;
    }
  }
}

// Entry e's internal key and value to keys[key_off[e]..] and vals[val_off[e]..].
__global__ void k_emit(Planes K, const u64* vptr, const u32* vlen, u64 n, const u64* key_off, const u64* val_off,
                       u8* keys, u8* vals) {
  GRID_STRIDE(e, n) {
    const u32 ul = K.klen[e];
    u8* kd = keys + key_off[e];
    for (u32 i = 0; i < ul; i++) kd[i] = ukey_byte(K, e, i);
    const u64 tag = K.tag[e];
    for (u32 i = 0; i < 8; i++) kd[ul + i] = (u8)(tag >> (8 * i));
    const GMEM u8* vs = gbytes(vptr[e]);
    u8* vd = vals + val_off[e];
    const u32 vl = vlen[e];
    for (u32 i = 0; i < vl; i++) vd[i] = vs[i];
  }
}

u32 grid_for(u64 n, u32 per_block) {
  u64 g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return (u32)g;
}

}  // namespace

void launch_entry_out_sizes(KTimer& T, Planes K, u64 n, u32* ksz, u32* vlen) {
  KLAUNCH(T, "entry_out_sizes", k_entry_out_sizes, grid_for(n + 1, 256), 256, 0, K, n, ksz, vlen);
}

void launch_emit(KTimer& T, Planes K, const u64* vptr, const u32* vlen, u64 n, const u64* key_off, const u64* val_off,
                 u8* keys, u8* vals) {
  if (!n) return;
  KLAUNCH(T, "emit", k_emit, grid_for(n, 256), 256, 0, K, vptr, vlen, n, key_off, val_off, keys, vals);
}

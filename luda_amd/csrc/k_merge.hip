// k_merge.hip — k-way merge of the compaction's input streams in
// InternalKeyComparator order (db/dbformat.cc:51-67), and the compaction's
// per-entry selection fused into the last merge round.
//
// MergingIterator (table/merger.cc:148-161) emits the smallest head and, on a
// tie, the child with the lower index.  For sorted children that output is
// the stable merge of the children in index order, which a tree of pairwise
// stable merges of adjacent runs reproduces: each round merges runs
// (0,1), (2,3), ... into the same index range, the left run winning ties.
// Each merge is merge-path partitioned.  k_merge_tiles: a workgroup takes a
// tile of outputs, finds where its two diagonals cross the merge path with a
// 128-way search per diagonal, stages the keys of its stretch of both runs
// in LDS and merges from there.  The first round reads the children in place
// (no starting permutation) and checks, from the staged keys, that every
// child is strictly increasing.  k_merge_select is the last round with the
// rest of DoCompactionWork's per-entry loop (db/db_impl.cc:1090-1157) in its
// epilogue: the drop rule against each output's predecessor, the
// grandparent count ShouldStopBefore walks, the kept entries' positions by a
// single-pass look-back scan across tiles, and the kept stream written
// straight from LDS — one pass over the key planes instead of merge, drop,
// keep scan and gather.  k_merge (key planes wider than 8 words): a thread
// binary-searches its diagonal in global memory and emits kItems.
#include <algorithm>

#include "engine.h"
#include "scan.h"
#include "select.h"

namespace {

constexpr u32 kItems = 8;

__device__ __forceinline__ bool key_less(const Planes& K, u32 a, u32 b) { return ikey_cmp(K, a, b) < 0; }

__global__ __launch_bounds__(256) void k_merge(const MergePair* pairs, u32 npairs, u32 total_tiles, const u32* in,
                                               u32* out, Planes K) {
  GRID_STRIDE(t, total_tiles) {
    u32 lo = 0, hi = npairs;
    while (hi - lo > 1) {
      const u32 mid = (lo + hi) >> 1;
      if (pairs[mid].tile0 <= t) lo = mid;
      else hi = mid;
    }
    const MergePair pr = pairs[lo];
    const u32 na = pr.a1 - pr.a0, nb = pr.b1 - pr.a1;
    const u64 d = (t - pr.tile0) * (u64)kItems;
    if (d >= (u64)na + nb) continue;
    const u32* A = in + pr.a0;
    const u32* B = in + pr.a1;
    u32 l = d > nb ? (u32)(d - nb) : 0u;
    u32 h = d < na ? (u32)d : na;
    while (l < h) {
      const u32 m = (l + h) >> 1;
      if (!key_less(K, B[d - m - 1], A[m])) l = m + 1;
      else h = m;
    }
    u32 i = l, j = (u32)(d - l);
    const u64 end = min(d + kItems, (u64)na + nb);
    for (u64 o = d; o < end; o++) {
      const bool takeA = j >= nb || (i < na && !key_less(K, B[j], A[i]));
      out[pr.a0 + o] = takeA ? A[i++] : B[j++];
    }
  }
}

// ---------------------------------------------------------------------------
// Tiled merge.

template <u32 NW>
struct MergeTile {
  static constexpr u32 kTile = NW <= 2 ? 1024 : (NW <= 4 ? 512 : 256);  // outputs per workgroup
};

// Keys of the tile's stretch of the two runs, in LDS: A part at [0, na),
// B part at [na, na + nb); slots kTile and kTile + 1 hold the items just
// before the two stretches (the fused round's candidates for the output
// before the tile).
template <u32 NW>
struct TileKeys {
  u64 w[NW][MergeTile<NW>::kTile + 2];
  u64 tag[MergeTile<NW>::kTile + 2];
  u32 len[MergeTile<NW>::kTile + 2];
  u32 idx[MergeTile<NW>::kTile + 2];
};

template <u32 NW>
__device__ __forceinline__ bool lds_less(const TileKeys<NW>& t, u32 x, u32 y) {
#pragma unroll
  for (u32 w = 0; w < NW; w++) {
    const u64 a = t.w[w][x], b = t.w[w][y];
    if (a != b) return a < b;
  }
  if (t.len[x] != t.len[y]) return t.len[x] < t.len[y];
  return t.tag[x] > t.tag[y];
}

template <u32 NW>
__device__ __forceinline__ bool lds_ukey_eq(const TileKeys<NW>& t, u32 x, u32 y) {
  if (t.len[x] != t.len[y]) return false;
#pragma unroll
  for (u32 w = 0; w < NW; w++)
    if (t.w[w][x] != t.w[w][y]) return false;
  return true;
}

// ukey_cmp_fixed / ikey_cmp_fixed (select.h) for a staged key.
template <u32 NW>
__device__ __forceinline__ int lds_ukey_cmp_fixed(const TileKeys<NW>& t, u32 x, const u64* fw, u32 W, u32 flen) {
#pragma unroll
  for (u32 w = 0; w < NW; w++) {
    if (w >= W) break;
    const u64 a = t.w[w][x], b = fw[w];
    if (a != b) return a < b ? -1 : 1;
  }
  const u32 la = t.len[x];
  if (la != flen) return la < flen ? -1 : 1;
  return 0;
}

// IsBaseLevelForKey (is_base_level_for_key in select.h) for a staged key.
template <u32 NW>
__device__ __forceinline__ bool lds_is_base_level(const TileKeys<NW>& t, u32 x, u32 W, const DropParams& dp) {
  for (u32 L = dp.first_level; L < 7; L++) {
    const u32 b = dp.lvl_begin[L], en = dp.lvl_end[L];
    if (b >= en) continue;
    u32 lo = b, hi = en;
    while (lo < hi) {
      const u32 m = (lo + hi) >> 1;
      if (lds_ukey_cmp_fixed(t, x, dp.deep_words + (2ull * m + 1) * W, W, dp.deep_len[2 * m + 1]) > 0) lo = m + 1;
      else hi = m;
    }
    if (lo < en && lds_ukey_cmp_fixed(t, x, dp.deep_words + 2ull * lo * W, W, dp.deep_len[2 * lo]) >= 0) return false;
  }
  return true;
}

// grandparents_below (select.h) for a staged key.
template <u32 NW>
__device__ __forceinline__ u32 lds_grandparents_below(const TileKeys<NW>& t, u32 x, u32 W, const DropParams& dp) {
  u32 lo = 0, hi = dp.ngp;
  while (lo < hi) {
    const u32 m = (lo + hi) >> 1;
    int c = lds_ukey_cmp_fixed(t, x, dp.gp_words + (u64)m * W, W, dp.gp_len[m]);
    if (!c && t.tag[x] != dp.gp_tag[m]) c = t.tag[x] > dp.gp_tag[m] ? -1 : 1;
    if (c > 0) lo = m + 1;
    else hi = m;
  }
  return lo;
}

template <u32 NW>
__device__ __forceinline__ void stage_key(TileKeys<NW>& tk, const Planes& K, u32 slot, u32 e) {
  tk.idx[slot] = e;
#pragma unroll
  for (u32 w = 0; w < NW; w++) tk.w[w][slot] = w < K.W ? K.P[(u64)w * K.stride + e] : 0;
  tk.len[slot] = K.klen[e];
  tk.tag[slot] = K.tag[e];
}

// A run of the index array being merged: entry at position i of the run.
// The first round has no array yet: the children are read in place.
struct Run {
  const u32* p;
  u32 base;
  __device__ __forceinline__ u32 operator[](u32 i) const { return p ? p[base + i] : base + i; }
};

// Where diagonal d of the merge of A (na) and B (nb) crosses the merge path:
// the number of A items among the first d outputs (A wins ties).  One
// wavefront per diagonal, 64 probes per round (a ballot narrows the range
// 65-fold), so a diagonal of a 256 MiB job takes three or four dependent
// loads and no barrier.
__device__ __forceinline__ u32 wave_split(const Planes& K, const Run& A, const Run& B, u32 na, u32 nb, u64 d) {
  const u32 lane = threadIdx.x & 63;
  u32 lo = d > nb ? (u32)(d - nb) : 0u;  // the answer is in [lo, hi]
  u32 hi = d < na ? (u32)d : na;
  while (lo < hi) {
    // probe m in [lo, hi): the answer is > m iff B[d-m-1] does not precede A[m]
    const u32 span = hi - lo;
    const u32 m = lo + (u32)(((u64)span * lane) / 64);
    const bool probe = m < hi && (lane == 0 || m != lo + (u32)(((u64)span * (lane - 1)) / 64));
    const bool gt = probe && ikey_cmp(K, B[(u32)(d - m - 1)], A[m]) >= 0;
    const u64 bg = __ballot(gt), bl = __ballot(probe && !gt);
    // probes are increasing in lane order and the predicate is monotone:
    // the answer lies after the last "greater" probe, at or before the first
    // "not greater" one
    if (bg) lo = (u32)__builtin_amdgcn_readlane((int)m, 63 - __builtin_clzll(bg)) + 1;
    if (bl) hi = (u32)__builtin_amdgcn_readlane((int)m, __builtin_ctzll(bl));
  }
  return lo;
}

// splits[t] = the A items before tile t's first output, for every tile of a
// round (tiles of pair p from pairs[p].tile0, `items` outputs each).
__global__ __launch_bounds__(256) void k_merge_splits(const MergePair* pairs, u32 npairs, u32 total_tiles,
                                                      const u32* in, Planes K, u32 items, u32* splits) {
  const u32 t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= total_tiles) return;
  u32 lo = 0, hi = npairs;
  while (hi - lo > 1) {
    const u32 mid = (lo + hi) >> 1;
    if (pairs[mid].tile0 <= t) lo = mid;
    else hi = mid;
  }
  const MergePair pr = pairs[lo];
  const u32 na = pr.a1 - pr.a0, nb = pr.b1 - pr.a1;
  const u64 d = std::min<u64>((u64)(t - pr.tile0) * items, (u64)na + nb);
  const u32 i = wave_split(K, Run{in, pr.a0}, Run{in, pr.a1}, na, nb, d);
  if ((threadIdx.x & 63) == 0) splits[t] = i;
}

// Tile t's split pair: its own split and the next tile's (or the end of its
// pair's A run).
__device__ __forceinline__ void tile_splits(const u32* splits, u32 t, u32 total_tiles, const MergePair& pr,
                                            u32 tiles_in_pair_end, u32* i0, u32* i1) {
  *i0 = splits[t];
  *i1 = t + 1 < tiles_in_pair_end ? splits[t + 1] : pr.a1 - pr.a0;
  (void)total_tiles;
}

// Stages the tile's stretches of A ([i0, i0 + ja)) and B ([b0, b0 + jb));
// with check, also verifies that each stretch continues its run in strictly
// increasing order (the run's item before the stretch included), which is
// what makes the stable merge equal MergingIterator's output; an unsorted
// child is left to the CPU path.  Ends with a barrier.
template <u32 NW>
__device__ __forceinline__ void stage_tile(TileKeys<NW>& tk, const Planes& K, const Run& A, const Run& B, u32 i0,
                                           u32 ja, u32 b0, u32 jb, bool check, DevStatus* st) {
  const u32 tid = threadIdx.x;
  for (u32 x = tid; x < ja + jb; x += 256) stage_key(tk, K, x, x < ja ? A[i0 + x] : B[b0 + (x - ja)]);
  __syncthreads();
  if (check) {
    bool bad = false;
    for (u32 x = tid; x < ja + jb; x += 256) {
      const bool inA = x < ja;
      const u32 first = inA ? 0u : ja;       // first staged item of this stretch
      const u32 pos = inA ? i0 + x : b0 + (x - ja);  // position in its run
      if (pos == 0) continue;
      if (x > first) bad |= !lds_less(tk, x - 1, x);
      else bad |= ikey_cmp(K, inA ? A[pos - 1] : B[pos - 1], tk.idx[x]) >= 0;
    }
    if (bad) set_err(st, ERR_UNSUPPORTED, SITE_UNSORTED);
  }
}

// This thread's kPer outputs [dl, dl + kPer) of a staged tile (ja A items,
// jb B items): a diagonal search in LDS, then a short merge; emit(o, x) gets
// each output o with its staged item x.
template <u32 NW, u32 kPer, typename Emit>
__device__ __forceinline__ void merge_staged(const TileKeys<NW>& tk, u32 ja, u32 jb, Emit emit) {
  const u32 dl = threadIdx.x * kPer;
  if (dl >= ja + jb) return;
  u32 l = dl > jb ? dl - jb : 0u, h = dl < ja ? dl : ja;
  while (l < h) {
    const u32 m = (l + h) >> 1;
    if (!lds_less(tk, ja + (dl - m - 1), m)) l = m + 1;
    else h = m;
  }
  u32 i = l, j = dl - l;
  const u32 end = std::min<u32>(dl + kPer, ja + jb);
  for (u32 o = dl; o < end; o++) {
    const bool takeA = j >= jb || (i < ja && !lds_less(tk, ja + j, i));
    emit(o, takeA ? i++ : ja + j++);
  }
}

template <u32 NW>
__global__ __launch_bounds__(256) void k_merge_tiles(const MergePair* pairs, u32 npairs, u32 total_tiles, const u32* in,
                                                     u32* out, Planes K, u32 check, const u32* splits, DevStatus* st) {
  constexpr u32 kTile = MergeTile<NW>::kTile;
  constexpr u32 kPer = kTile / 256;
  __shared__ TileKeys<NW> tk;
  for (u32 t = blockIdx.x; t < total_tiles; t += gridDim.x) {
    u32 lo = 0, hi = npairs;
    while (hi - lo > 1) {
      const u32 mid = (lo + hi) >> 1;
      if (pairs[mid].tile0 <= t) lo = mid;
      else hi = mid;
    }
    const MergePair pr = pairs[lo];
    const u32 na = pr.a1 - pr.a0, nb = pr.b1 - pr.a1;
    const u64 d0 = (u64)(t - pr.tile0) * kTile;
    if (d0 >= (u64)na + nb) continue;  // uniform over the workgroup
    const u64 d1 = std::min<u64>(d0 + kTile, (u64)na + nb);
    const Run A{in, pr.a0}, B{in, pr.a1};
    const u32 pair_end = pr.tile0 + (u32)(((u64)na + nb + kTile - 1) / kTile);
    u32 i0, i1;
    tile_splits(splits, t, total_tiles, pr, pair_end, &i0, &i1);
    const u32 ja = i1 - i0;                                   // A items in the tile
    const u32 jb = (u32)(d1 - d0) - ja;                       // B items in the tile
    stage_tile(tk, K, A, B, i0, ja, (u32)(d0 - i0), jb, check != 0, st);
    u32* dst = out + pr.a0 + d0;
    merge_staged<NW, kPer>(tk, ja, jb, [&](u32 o, u32 x) { dst[o] = tk.idx[x]; });
    __syncthreads();
  }
}

// The last merge round (one pair, one workgroup per tile, tiles in launch
// order so that the look-back only ever waits on tiles dispatched earlier)
// with the selection of DoCompactionWork's loop:
//   drop(p) = parsed(p) && ((parsed(p-1) && ukey(p-1) == ukey(p) && seq(p-1) <= S)
//                           || (type(p) == deletion && seq(p) <= S && IsBaseLevelForKey(ukey(p))))
// (db/db_impl.cc:1113-1147: last_sequence_for_key is the predecessor's
// sequence when the predecessor parsed and has the same user key; a key
// whose type is not a deletion or a value does not parse, is kept and resets
// that state), then the kept entries' output positions and the kept stream.
struct SelectArgs {
  Planes K;            // decoded entries
  const u64* vptr;
  const u32* vlen;
  DropParams dp;
  Kept out;            // the kept stream
  u32* gcnt;           // per merged position, with grandparents only
  u32* kpos;           // per merged position: kept entries before it (with grandparents only)
  u32* slot;           // crossing slots (cleared before the launch)
  ScanArgs sa;         // look-back words of the kept-count scan
};

template <u32 NW>
__global__ __launch_bounds__(256) void k_merge_select(MergePair pr, const u32* in, u32 check, const u32* splits,
                                                      SelectArgs s) {
  constexpr u32 kTile = MergeTile<NW>::kTile;
  constexpr u32 kPer = kTile / 256;
  __shared__ TileKeys<NW> tk;
  __shared__ u16 ord[kTile];   // output o of the tile -> its staged item
  __shared__ u16 kx[kTile];    // kept output r of the tile -> its staged item
  __shared__ u32 lastg[256];   // grandparent count of each thread's last output
  __shared__ u32 wsum[4];
  __shared__ u32 s_pre, s_ticket;
  const Planes& K = s.K;
  const DropParams& dp = s.dp;
  DevStatus* st = s.sa.st;
  const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const u32 t = lbscan::take_ticket(s.sa, &s_ticket);  // the tile, in the order workgroups start
  const u32 na = pr.a1 - pr.a0, nb = pr.b1 - pr.a1;
  const u64 n = (u64)na + nb;
  const u64 tiles = (n + kTile - 1) / kTile;
  const u64 d0 = (u64)t * kTile;
  const u64 d1 = std::min<u64>(d0 + kTile, n);
  const Run A{in, pr.a0}, B{in, pr.a1};
  u32 i0, i1;
  tile_splits(splits, t, (u32)tiles, pr, (u32)tiles, &i0, &i1);
  const u32 ja = i1 - i0, jb = (u32)(d1 - d0) - ja, b0 = (u32)(d0 - i0);
  const bool has_prev = d0 > 0;
  // the candidates for the output before the tile, staged beside it
  if (tid == 254 && i0 > 0) stage_key(tk, K, kTile, A[i0 - 1]);
  if (tid == 255 && b0 > 0) stage_key(tk, K, kTile + 1, B[b0 - 1]);
  stage_tile(tk, K, A, B, i0, ja, b0, jb, check != 0, st);
  merge_staged<NW, kPer>(tk, ja, jb, [&](u32 o, u32 x) { ord[o] = (u16)x; });
  __syncthreads();
  // the output before the tile: the later of A[i0 - 1] and B[b0 - 1] in
  // merge order (A precedes B on a tie)
  const u32 prev = i0 == 0 ? kTile + 1 : (b0 == 0 ? kTile : (lds_less(tk, kTile + 1, kTile) ? kTile : kTile + 1));
  // drop rule and grandparent count of this thread's outputs, from LDS
  const u32 W = K.W;
  const u32 nout = (u32)(d1 - d0);
  const u32 o0 = tid * kPer;
  u32 keepm = 0, cnt = 0;
  u32 g[kPer];
#pragma unroll
  for (u32 k = 0; k < kPer; k++) {
    const u32 o = o0 + k;
    g[k] = 0;
    if (o >= nout) continue;
    const u32 x = ord[o];
    const u64 tag = tk.tag[x];
    const u32 type = (u32)(tag & 0xff);
    bool drop = false;
    if (type <= 1) {
      if (o > 0 || has_prev) {
        const u32 px = o > 0 ? ord[o - 1] : prev;
        const u64 tp = tk.tag[px];
        if ((tp & 0xff) <= 1 && (tp >> 8) <= dp.S && lds_ukey_eq(tk, px, x)) drop = true;
      }
      if (!drop && type == 0 && (tag >> 8) <= dp.S && lds_is_base_level(tk, x, W, dp)) drop = true;
    }
    if (!drop) {
      keepm |= 1u << k;
      cnt++;
    }
    if (dp.ngp) g[k] = lds_grandparents_below(tk, x, W, dp);
  }
  if (dp.ngp) {
    u32 last = 0;
#pragma unroll
    for (u32 k = 0; k < kPer; k++)
      if (o0 + k < nout) last = g[k];
    lastg[tid] = last;
  }
  // kept entries before this thread's: wave scan, wave sums, then the
  // tiles before this one by look-back
  const u32 incl = wave_incl_sum(cnt);
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  u32 r = incl - cnt;
  for (u32 w = 0; w < wave; w++) r += wsum[w];
  const u32 tile_kept = wsum[0] + wsum[1] + wsum[2] + wsum[3];
#pragma unroll
  for (u32 k = 0; k < kPer; k++)
    if (keepm & (1u << k)) kx[r++] = ord[o0 + k];
  u32 gp_before = 0;  // grandparent count of the output before this thread's first
  if (dp.ngp && o0 < nout) {
    if (tid > 0) gp_before = lastg[tid - 1];
    else if (has_prev) gp_before = lds_grandparents_below(tk, prev, W, dp);
  }
  if (wave == 0) {
    const u32 ex = lbscan::publish<u32>(s.sa, t, tile_kept);
    if (lane == 0) {
      s_pre = ex;
      if (t + 1 == tiles) st->n_kept = ex + tile_kept;
    }
  }
  __syncthreads();
  const u32 pre = s_pre;
  if (dp.ngp && o0 < nout) {
    // per merged position: grandparent count, kept entries before it, and
    // the crossing slots
    u32 rk = pre + r - cnt;
#pragma unroll
    for (u32 k = 0; k < kPer; k++) {
      const u32 o = o0 + k;
      if (o >= nout) break;
      const u64 p = d0 + o;
      s.gcnt[p] = g[k];
      s.kpos[p] = rk;
      rk += (keepm >> k) & 1u;
      const u32 pg = k > 0 ? g[k - 1] : gp_before;
      if ((o > 0 || has_prev) && g[k] > pg) s.slot[g[k]] = (u32)p;
    }
  }
  // the tile's kept entries, one per thread in order: coalesced stores
  const Kept& out = s.out;
  for (u32 j = tid; j < tile_kept; j += 256) {
    const u32 x = kx[j];
    const u64 q = pre + j;
    const u32 e = tk.idx[x];
#pragma unroll
    for (u32 w = 0; w < NW; w++)
      if (w < out.K.W) out.K.P[(u64)w * out.K.stride + q] = tk.w[w][x];
    out.K.klen[q] = tk.len[x];
    out.K.tag[q] = tk.tag[x];
    out.vptr[q] = s.vptr[e];
    out.vlen[q] = s.vlen[e];
    out.forced[q] = 0;
  }
}

// Entries of each child (child c = [cs[c], cs[c+1])) must be strictly
// increasing: that is what makes the stable merge equal MergingIterator's
// output.  A table that is not sorted is left to the CPU path.
__global__ void k_check_sorted(Planes K, u64 n, const u32* cs, u32 nchild, u32* iota, DevStatus* st) {
  GRID_STRIDE(e, n) {
    iota[e] = (u32)e;  // the merge's starting permutation
    if (e == 0) continue;
    u32 lo = 0, hi = nchild;
    while (hi - lo > 1) {
      const u32 mid = (lo + hi) >> 1;
      if (cs[mid] <= e) lo = mid;
      else hi = mid;
    }
    if (cs[lo] == e) continue;
    if (ikey_cmp(K, e - 1, e) >= 0) set_err(st, ERR_UNSUPPORTED, SITE_UNSORTED);
  }
}

u32 grid_for(u64 n, u32 per_block) {
  u64 g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return (u32)g;
}

}  // namespace

u32 merge_items_per_tile(u32 W) {
  if (W <= 2) return MergeTile<2>::kTile;
  if (W <= 4) return MergeTile<4>::kTile;
  if (W <= 8) return MergeTile<8>::kTile;
  return kItems;
}

void launch_merge(KTimer& T, const MergePair* pairs, u32 npairs, u32 total_tiles, const u32* in, u32* out, Planes K,
                  bool check, u32* splits, DevStatus* st) {
  if (!total_tiles) return;
  const u32 g = std::min<u32>(total_tiles, 4096);
  const u32 c = check ? 1u : 0u;
  if (K.W <= 8)
    KLAUNCH(T, "merge_splits", k_merge_splits, (total_tiles + 3) / 4, 256, 0, pairs, npairs, total_tiles, in, K,
            merge_items_per_tile(K.W), splits);
  if (K.W <= 2)
    KLAUNCH(T, "merge", k_merge_tiles<2>, g, 256, 0, pairs, npairs, total_tiles, in, out, K, c, (const u32*)splits, st);
  else if (K.W <= 4)
    KLAUNCH(T, "merge", k_merge_tiles<4>, g, 256, 0, pairs, npairs, total_tiles, in, out, K, c, (const u32*)splits, st);
  else if (K.W <= 8)
    KLAUNCH(T, "merge", k_merge_tiles<8>, g, 256, 0, pairs, npairs, total_tiles, in, out, K, c, (const u32*)splits, st);
  else KLAUNCH(T, "merge", k_merge, grid_for(total_tiles, 256), 256, 0, pairs, npairs, total_tiles, in, out, K);
}

void launch_merge_select(KTimer& T, ScanState& S, const MergePair* d_pair, MergePair pr, const u32* in, bool check,
                         Planes K, const u64* vptr, const u32* vlen, const DropParams& dp, Kept out, u32* gcnt,
                         u32* kpos, u32* slot, u32* splits, DevStatus* st) {
  const u64 n = (u64)pr.b1 - pr.a0;
  if (!n) return;
  const u64 tiles = (n + merge_items_per_tile(K.W) - 1) / merge_items_per_tile(K.W);
  if (tiles > S.tiles) throw_scan_capacity(tiles, S.tiles);
  SelectArgs a;
  a.K = K;
  a.vptr = vptr;
  a.vlen = vlen;
  a.dp = dp;
  a.out = out;
  a.gcnt = gcnt;
  a.kpos = kpos;
  a.slot = slot;
  a.sa.words = S.words;
  a.sa.ticket = S.tickets ? S.ticket : nullptr;
  a.sa.epoch = scan_next_epoch(S, T.s);
  a.sa.cap = 0;
  a.sa.need = nullptr;
  a.sa.st = st;
  const u32 c = check ? 1u : 0u;
  if (K.W > 8) launch_refused("merge_select (key planes wider than 8 words)", hipErrorInvalidValue);
  KLAUNCH(T, "merge_splits", k_merge_splits, (u32)((tiles + 3) / 4), 256, 0, d_pair, 1u, (u32)tiles, in, K,
          merge_items_per_tile(K.W), splits);
  const u32* sp = splits;
  if (K.W <= 2) KLAUNCH(T, "merge_select", k_merge_select<2>, (u32)tiles, 256, 0, pr, in, c, sp, a);
  else if (K.W <= 4) KLAUNCH(T, "merge_select", k_merge_select<4>, (u32)tiles, 256, 0, pr, in, c, sp, a);
  else KLAUNCH(T, "merge_select", k_merge_select<8>, (u32)tiles, 256, 0, pr, in, c, sp, a);
}
void launch_check_sorted(KTimer& T, Planes K, u64 n, const u32* child_start, u32 nchild, u32* iota, DevStatus* st) {
  if (!n) return;
  KLAUNCH(T, "check_sorted", k_check_sorted, grid_for(n, 256), 256, 0, K, n, child_start, nchild, iota, st);
}

// k_merge.hip — k-way merge of the compaction's input streams in
// InternalKeyComparator order (db/dbformat.cc:51-67).
//
// MergingIterator (table/merger.cc:148-161) emits the smallest head and, on a
// tie, the child with the lower index.  For sorted children that output is
// the stable merge of the children in index order, which a tree of pairwise
// stable merges of adjacent runs reproduces: each round merges runs
// (0,1), (2,3), ... into the same index range, the left run winning ties.
// Each merge is merge-path partitioned.  k_merge_tiles: a workgroup takes a
// tile of outputs, finds where its two diagonals cross the merge path with a
// 128-way search per diagonal, stages the keys of its stretch of both runs
// in LDS and merges from there.  k_merge (key planes wider than 8 words): a
// thread binary-searches its diagonal in global memory and emits kItems.
#include <algorithm>

#include "engine.h"

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
// B part at [na, na + nb).
template <u32 NW>
struct TileKeys {
  u64 w[NW][MergeTile<NW>::kTile];
  u64 tag[MergeTile<NW>::kTile];
  u32 len[MergeTile<NW>::kTile];
  u32 idx[MergeTile<NW>::kTile];
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

// Where diagonals d0 and d1 of the merge of A (na) and B (nb) cross the
// merge path: the number of A items among the first d outputs (A wins ties).
// Threads 0-127 search d0 and 128-255 search d1, 128 probes per round each,
// until both are found; every barrier is reached by the whole workgroup.
// box: 2 x 2 words of LDS.
__device__ __forceinline__ void coop_splits(const Planes& K, const u32* A, const u32* B, u32 na, u32 nb, u64 d0,
                                            u64 d1, u32 (*box)[2], u32* i0, u32* i1) {
  const u32 half = threadIdx.x >> 7, sub = threadIdx.x & 127;
  const u64 d = half ? d1 : d0;
  u32 lo = d > nb ? (u32)(d - nb) : 0u;  // the answer is in [lo, hi]
  u32 hi = d < na ? (u32)d : na;
  for (;;) {
    // probe m in [lo, hi): the answer is > m iff B[d-m-1] does not precede A[m]
    const u32 span = hi - lo;
    const u32 m = lo + (u32)(((u64)span * (sub + 1)) / 129);
    const bool probe = lo < hi && m < hi && (sub == 0 || m != lo + (u32)(((u64)span * sub) / 129));
    const bool gt = probe && ikey_cmp(K, B[d - m - 1], A[m]) >= 0;
    if (sub == 0) {
      box[half][0] = lo;
      box[half][1] = hi;
    }
    __syncthreads();
    if (probe) {
      if (gt) atomicMax(&box[half][0], m + 1);
      else atomicMin(&box[half][1], m);
    }
    __syncthreads();
    lo = box[half][0];
    hi = box[half][1];
    const bool fin = box[0][0] >= box[0][1] && box[1][0] >= box[1][1];
    __syncthreads();
    if (fin) break;
  }
  *i0 = box[0][0];
  *i1 = box[1][0];
}

template <u32 NW>
__global__ __launch_bounds__(256) void k_merge_tiles(const MergePair* pairs, u32 npairs, u32 total_tiles, const u32* in,
                                                     u32* out, Planes K) {
  constexpr u32 kTile = MergeTile<NW>::kTile;
  constexpr u32 kPer = kTile / 256;
  __shared__ TileKeys<NW> tk;
  __shared__ u32 box[2][2];
  const u32 tid = threadIdx.x;
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
    const u32* A = in + pr.a0;
    const u32* B = in + pr.a1;
    u32 i0, i1;
    coop_splits(K, A, B, na, nb, d0, d1, box, &i0, &i1);
    const u32 ja = i1 - i0;                                   // A items in the tile
    const u32 jb = (u32)(d1 - d0) - ja;                       // B items in the tile
    const u32 b0 = (u32)(d0 - i0);
    // stage keys: A stretch then B stretch
    for (u32 x = tid; x < ja + jb; x += 256) {
      const u32 e = x < ja ? A[i0 + x] : B[b0 + (x - ja)];
      tk.idx[x] = e;
#pragma unroll
      for (u32 w = 0; w < NW; w++) tk.w[w][x] = w < K.W ? K.P[(u64)w * K.stride + e] : 0;
      tk.len[x] = K.klen[e];
      tk.tag[x] = K.tag[e];
    }
    __syncthreads();
    // this thread's kPer outputs: diagonal search in LDS, then a short merge
    const u32 dl = tid * kPer;
    if (dl < ja + jb) {
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
        out[pr.a0 + d0 + o] = takeA ? tk.idx[i++] : tk.idx[ja + j++];
      }
    }
    __syncthreads();
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

void launch_merge(KTimer& T, const MergePair* pairs, u32 npairs, u32 total_tiles, const u32* in, u32* out, Planes K) {
  if (!total_tiles) return;
  const u32 g = std::min<u32>(total_tiles, 4096);
  if (K.W <= 2) KLAUNCH(T, "merge", k_merge_tiles<2>, g, 256, 0, pairs, npairs, total_tiles, in, out, K);
  else if (K.W <= 4) KLAUNCH(T, "merge", k_merge_tiles<4>, g, 256, 0, pairs, npairs, total_tiles, in, out, K);
  else if (K.W <= 8) KLAUNCH(T, "merge", k_merge_tiles<8>, g, 256, 0, pairs, npairs, total_tiles, in, out, K);
  else KLAUNCH(T, "merge", k_merge, grid_for(total_tiles, 256), 256, 0, pairs, npairs, total_tiles, in, out, K);
}

void launch_check_sorted(KTimer& T, Planes K, u64 n, const u32* child_start, u32 nchild, u32* iota, DevStatus* st) {
  if (!n) return;
  KLAUNCH(T, "check_sorted", k_check_sorted, grid_for(n, 256), 256, 0, K, n, child_start, nchild, iota, st);
}

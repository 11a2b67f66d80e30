// k_merge.hip — k-way merge of the compaction's input streams in
// InternalKeyComparator order (db/dbformat.cc:51-67).
//
// MergingIterator (table/merger.cc:148-161) emits the smallest head and, on a
// tie, the child with the lower index.  For sorted children that output is
// the stable merge of the children in index order, which a tree of pairwise
// stable merges of adjacent runs reproduces: each round merges runs
// (0,1), (2,3), ... into the same index range, the left run winning ties.
// Each merge is merge-path partitioned: a thread binary-searches its output
// diagonal and then emits kItems outputs.
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

// Entries of each child (child c = [cs[c], cs[c+1])) must be strictly
// increasing: that is what makes the stable merge equal MergingIterator's
// output.  A table that is not sorted is left to the CPU path.
__global__ void k_check_sorted(Planes K, u64 n, const u32* cs, u32 nchild, DevStatus* st) {
  GRID_STRIDE(e, n) {
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

u32 merge_items_per_tile() { return kItems; }

void launch_merge(KTimer& T, const MergePair* pairs, u32 npairs, u32 total_tiles, const u32* in, u32* out, Planes K) {
  if (!total_tiles) return;
  KLAUNCH(T, "merge", k_merge, grid_for(total_tiles, 256), 256, 0, pairs, npairs, total_tiles, in, out, K);
}

void launch_check_sorted(KTimer& T, Planes K, u64 n, const u32* child_start, u32 nchild, DevStatus* st) {
  if (n < 2) return;
  KLAUNCH(T, "check_sorted", k_check_sorted, grid_for(n, 256), 256, 0, K, n, child_start, nchild, st);
}

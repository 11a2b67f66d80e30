// select.h — per-entry predicates of DBImpl::DoCompactionWork's loop, shared
// by the units that evaluate them (k_select.hip, and the fused final merge
// round of k_merge.hip): IsBaseLevelForKey over the deeper levels' sorted
// file ranges, and the count of grandparents whose largest key is below an
// entry (what ShouldStopBefore walks past).
#pragma once

#include "engine.h"

namespace {

// Internal-key compare of plane entry e against a stored key.
__device__ __forceinline__ int ikey_cmp_fixed(const Planes& K, u64 e, const u64* words, u32 len, u64 tag) {
  int c = ukey_cmp_fixed(K, e, words, len);
  if (c) return c;
  const u64 te = K.tag[e];
  if (te != tag) return te > tag ? -1 : 1;
  return 0;
}

// Compaction::IsBaseLevelForKey (db/version_set.cc:1535-1554): no file of a
// level >= level + 2 has smallest <= key <= largest.  The reference walks
// monotone per-level pointers; for a sorted, disjoint level that equals a
// binary search for the first file whose largest key is >= the key.
__device__ __forceinline__ bool is_base_level_for_key(const Planes& K, u64 e, const DropParams& dp) {
  for (u32 L = dp.first_level; L < 7; L++) {
    const u32 b = dp.lvl_begin[L], en = dp.lvl_end[L];
    if (b >= en) continue;
    u32 lo = b, hi = en;
    while (lo < hi) {
      const u32 m = (lo + hi) >> 1;
      if (ukey_cmp_fixed(K, e, dp.deep_words + (2ull * m + 1) * K.W, dp.deep_len[2 * m + 1]) > 0) lo = m + 1;
      else hi = m;
    }
    if (lo < en && ukey_cmp_fixed(K, e, dp.deep_words + 2ull * lo * K.W, dp.deep_len[2 * lo]) >= 0) return false;
  }
  return true;
}

// Grandparents whose largest internal key is below entry e: the count
// Compaction::ShouldStopBefore (db/version_set.cc:1556-1579) has advanced past
// when e arrives.
__device__ __forceinline__ u32 grandparents_below(const Planes& K, u64 e, const DropParams& dp) {
  u32 lo = 0, hi = dp.ngp;
  while (lo < hi) {
    const u32 m = (lo + hi) >> 1;
    if (ikey_cmp_fixed(K, e, dp.gp_words + (u64)m * K.W, dp.gp_len[m], dp.gp_tag[m]) > 0) lo = m + 1;
    else hi = m;
  }
  return lo;
}

// slot[g] of the grandparent crossings (g = 1..ngp): no crossing yet
constexpr u32 kNoEvent = 0xffffffffu;

}  // namespace

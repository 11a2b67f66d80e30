// cuts.h — the per-tile steps of the tiled output block cuts (k_cuts.hip):
// a data block ends once CurrentSizeEstimate() >= block_size after an Add
// (table/table_builder.cc:119-123) or before an entry that starts a table
// (db/db_impl.cc:1105-1111, the kept stream's `forced` flags).  Written once
// for the device and for a host build (plain C++, bounds-checked) that the
// CPU tests run against a one-entry-at-a-time restatement of that rule.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define CUTS_FN __host__ __device__ __forceinline__
#define CUTS_CHECK(c)
#else
#include <cassert>
#define CUTS_FN inline
#define CUTS_CHECK(c) assert(c)
#endif

namespace cuts {

constexpr uint32_t kTile = 4096;  // entries per tile
constexpr uint32_t kSpan = 1024;  // entries staged past the tile: a block of at most this many crosses its end
constexpr uint32_t kCand = 256;   // most entry points into a tile
constexpr uint32_t kStaged = kTile + kSpan;
constexpr uint32_t kUnresolved = 0xffffu;
static_assert(kStaged < kUnresolved, "block ends fit in 16 bits");
static_assert(kSpan < 4096 && kTile < (1u << 20), "a function entry packs exit | blocks << 12");

// A tile's staged entries.  P[x]: sznr bytes of entries [0, x); F[x]: forced
// entries among them; D[x]: szr - sznr, what entry x adds as a restart
// point; nx[x]: end of the block that would start at x.  (A tile whose bytes
// or surcharges do not fit these widths is left to the orbit kernels.)
struct Tile {
  uint32_t P[kStaged + 1];
  uint16_t F[kStaged + 1];
  uint16_t D[kStaged];
  uint16_t nx[kTile];
};

CUTS_FN uint32_t P_at(const Tile& L, uint32_t i) {
  CUTS_CHECK(i <= kStaged);
  return L.P[i];
}
CUTS_FN uint32_t F_at(const Tile& L, uint32_t i) {
  CUTS_CHECK(i <= kStaged);
  return L.F[i];
}
CUTS_FN uint32_t D_at(const Tile& L, uint32_t i) {
  CUTS_CHECK(i < kStaged);
  return L.D[i];
}
CUTS_FN uint32_t nx_at(const Tile& L, uint32_t i) {
  CUTS_CHECK(i < kTile);
  return L.nx[i];
}

// sznr bytes of entries [a, b)
CUTS_FN uint64_t bytes(const Tile& L, uint32_t a, uint32_t b) { return (uint64_t)(P_at(L, b) - P_at(L, a)); }

// CurrentSizeEstimate() of a block holding entries [x, y): entries, the
// restart points' surcharges, the restart array and its count.
CUTS_FN uint64_t estimate(const Tile& L, uint32_t x, uint32_t y, uint32_t R) {
  const uint32_t g = (y - x + R - 1) / R;
  uint64_t s = bytes(L, x, y) + 4ull * g + 4;
  for (uint32_t i = 0; i < g; i++) s += D_at(L, x + i * R);
  return s;
}

// Mean entries per block over the staged entries: where block_end starts
// its search.
CUTS_FN uint32_t mean_block(const Tile& L, uint32_t lim, uint64_t bs) {
  const uint64_t tb = P_at(L, lim);
  uint64_t e = tb ? (bs * lim) / tb : lim;
  if (e > lim) e = lim;
  return e ? (uint32_t)e : 1u;
}

// The end of the block that starts at entry x < lim: the first y with
// estimate(x, y) >= bs, cut first by a forced entry in (x, y) and last by the
// stream's end (stream_end: the staged entries reach it).  The estimate is at
// least bytes(x, y) + 8, so a galloping search (from est entries on) for the
// first y where that bound reaches bs bounds the end from above; a walk back
// then finds the exact end.  kUnresolved when the block runs past the staged
// entries.  *est_out: the block's estimate (its bytes without the trailer).
CUTS_FN uint32_t block_end(const Tile& L, uint32_t x, uint32_t lim, bool stream_end, uint32_t est, uint64_t bs,
                           uint32_t R, uint64_t* est_out) {
  CUTS_CHECK(x < lim && lim <= kStaged && est >= 1 && R >= 1);
  uint32_t limit = lim;
  const uint32_t f0 = F_at(L, x + 1);  // forced entries in [0, x]
  if (F_at(L, lim) > f0) {
    uint32_t lo = x + 1, hi = lim - 1;  // smallest e in [x + 1, lim) with forced[e]: F[e + 1] counts it
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (F_at(L, mid + 1) > f0) hi = mid;
      else lo = mid + 1;
    }
    limit = lo;
  }
  const uint32_t bx = P_at(L, x);
  auto reach = [&](uint32_t y) { return (uint64_t)(P_at(L, y) - bx) + 8 >= bs; };
  uint32_t lo = x + 1, hi = limit + 1;  // the first y that reaches lies in [lo, hi]; limit + 1: none
  const uint32_t y = x + est < limit ? x + est : limit;
  if (reach(y)) {
    hi = y;
    uint32_t d = 1;
    while (hi - lo >= d) {
      if (reach(hi - d)) {
        hi -= d;
        d <<= 1;
      } else {
        lo = hi - d + 1;
        break;
      }
    }
  } else {
    lo = y + 1;
    uint32_t d = 1;
    while (lo + d - 1 <= limit) {
      const uint32_t z = lo + d - 1;
      if (reach(z)) {
        hi = z;
        break;
      }
      lo = z + 1;
      d <<= 1;
    }
  }
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (reach(mid)) hi = mid;
    else lo = mid + 1;
  }
  uint32_t end = lo < limit ? lo : limit;
  uint64_t S = estimate(L, x, end, R);
  if (S >= bs) {
    // walk back while the block one entry shorter still reaches bs
    while (end > x + 1) {
      const uint32_t e = end - 1;
      const bool rs = (e - x) % R == 0;
      const uint64_t sz = bytes(L, e, e + 1) + (rs ? (uint64_t)D_at(L, e) + 4 : 0);
      if (S - sz < bs) break;
      S -= sz;
      end = e;
    }
    *est_out = S;
    return end;
  }
  *est_out = S;
  if (limit < lim) return limit;  // a forced entry starts the next block
  if (stream_end) return lim;
  return kUnresolved;
}

// How many entry points the chain can have into the tile: the first block
// start at or after its first entry is entry c only if the block before held
// entries [0, c) of the tile, whose bytes + 8 stay below bs, and no forced
// entry among them; c = te when the stream ends there.  0 when the tile
// cannot tell (a block may run past the staged entries).
CUTS_FN uint32_t entry_points(const Tile& L, uint32_t lim, uint32_t te, bool stream_end, uint64_t bs) {
  uint32_t lo = 0, hi = lim;  // smallest y in [0, lim] with bytes(0, y) + 8 >= bs, else lim
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((uint64_t)P_at(L, mid) + 8 >= bs) hi = mid;
    else lo = mid + 1;
  }
  if (lo == lim && (uint64_t)P_at(L, lim) + 8 < bs && !stream_end) return 0;
  uint32_t C = lo + 1;
  if (F_at(L, lim) > 0) {
    uint32_t flo = 0, fhi = lim - 1;  // the first forced entry
    while (flo < fhi) {
      const uint32_t mid = (flo + fhi) >> 1;
      if (F_at(L, mid + 1) > 0) fhi = mid;
      else flo = mid + 1;
    }
    if (flo + 1 < C) C = flo + 1;
  }
  if (te + 1 < C) C = te + 1;
  return C;
}

// The chain entering the tile at c: where it leaves (offset into the next
// tile) and the blocks it starts in this tile, packed exit | blocks << 12.
CUTS_FN uint32_t walk(const Tile& L, uint32_t c, uint32_t te) {
  uint32_t x = c, n = 0;
  while (x < te) {
    const uint32_t y = nx_at(L, x);
    CUTS_CHECK(y > x);
    x = y;
    n++;
  }
  CUTS_CHECK(x - te < 4096);
  return (x - te) | (n << 12);
}

}  // namespace cuts

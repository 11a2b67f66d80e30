// k_cuts.hip — where the table writer's data blocks end.  TableBuilder
// flushes a block once CurrentSizeEstimate() >= block_size after an Add
// (table/table_builder.cc:119-123), and DoCompactionWork ends the table (so
// also its block) before an entry that ShouldStopBefore cuts at
// (db/db_impl.cc:1105-1111).  The blocks are therefore a greedy chain over
// the kept entries: which entry ends a block depends on where it started.
//
// Three launches find the chain:
//  1. cut_tiles: per tile of kCutTile entries (staged in LDS with kCutSpan
//     more), every entry's block end as if a block started there, then, for
//     every entry point the chain can have into the tile (the first block
//     start at or after the tile's first entry: one of its first C entries),
//     the point where the chain leaves the tile and the blocks it made on the
//     way — the tile's function, C entries in global memory.
//  2. cut_chain: one workgroup composes the functions from the first tile on
//     (sixteen segments of tiles at once, then the segments in order), which
//     gives every tile its real entry point and the number of blocks before
//     it.
//  3. cut_emit: per tile, the real chain from the tile's entry point over the
//     block ends cut_tiles kept (bnext, bbytes), and each block's descriptor
//     (first entry, entries, bytes, table start).
// Shapes outside what the tiles hold (more than kCutCand entry points, a
// block longer than the staged span) set cut_fallback with ERR_OVERFLOW, and
// the host runs the job again with the orbit kernels.
#include "cuts.h"
#include "engine.h"

namespace {

using cuts::kStaged;
using cuts::kUnresolved;
constexpr u32 kCutThreads = 512;
constexpr u32 kStagePer = kStaged / kCutThreads;  // staged entries per thread (the prefix sum's run)
constexpr u32 kStartsPer = kCutTile / kCutThreads;
static_assert(kStaged % kCutThreads == 0 && kCutTile % kCutThreads == 0, "whole rows per thread");

// 49 KiB: three tiles per CU, every tile of a 256 MiB job in one round
struct CutLds {
  cuts::Tile t;
  u64 wsum[kCutThreads / 64];
  u32 fsum[kCutThreads / 64];
};

__device__ __forceinline__ void cut_fall_back(DevStatus* st) {
  st->cut_fallback = 1;
  set_err(st, ERR_OVERFLOW, SITE_BLOCK_ENTRIES);
}

// cut_tiles.  Tile t = blockIdx.x: stages the tile's entries (coalesced
// loads, then prefix sums over runs of kStagePer consecutive entries per
// thread), computes every start's block end and bytes (kept in bnext and
// bbytes, relative to the tile, for cut_emit) and the tile's function.
__global__ __launch_bounds__(kCutThreads) void k_cut_tiles(Kept k, const u32* m_dev, EncParams ep, EncBufs b,
                                                           CutBufs cb, DevStatus* st) {
  __shared__ CutLds S;
  __shared__ u32 s_skip;
  cuts::Tile& L = S.t;
  const u64 m = *m_dev;
  const u64 t = blockIdx.x;
  const u64 T0 = t * kCutTile;
  if (T0 >= m) return;
  const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // one read of the status for the whole workgroup: other tiles may set it
  // meanwhile, and every thread must take the same way to the barriers
  if (tid == 0) s_skip = st->err;
  __syncthreads();
  if (s_skip) return;
  const u32 te = (u32)min<u64>(kCutTile, m - T0);
  const u32 lim = (u32)min<u64>(kStaged, m - T0);
  {
    u32 r[kStagePer], nr[kStagePer], fo[kStagePer];
#pragma unroll
    for (u32 i = 0; i < kStagePer; i++) {
      const u32 x = i * kCutThreads + tid;
      const u64 q = T0 + (x < lim ? x : 0);
      r[i] = b.szr[q];
      nr[i] = b.sznr[q];
      fo[i] = k.forced[q];
    }
    bool bad = false, wide = false;
#pragma unroll
    for (u32 i = 0; i < kStagePer; i++) {
      const u32 x = i * kCutThreads + tid;
      if (x < lim) {
        bad |= r[i] >= (1u << 31) || nr[i] >= (1u << 31) || r[i] < nr[i];
        wide |= r[i] - nr[i] > 0xffffu;
        L.D[x] = (u16)(r[i] - nr[i]);
        L.P[x + 1] = nr[i];
        L.F[x + 1] = fo[i] ? 1 : 0;
      }
    }
    if (bad) set_err(st, ERR_UNSUPPORTED, SITE_BLOCK_ENTRIES);
    if (wide) cut_fall_back(st);
  }
  __syncthreads();
  u32 v[kStagePer], f[kStagePer];
  u64 tot = 0;
  u32 ftot = 0;
#pragma unroll
  for (u32 i = 0; i < kStagePer; i++) {
    const u32 x = kStagePer * tid + i;
    v[i] = x < lim ? L.P[x + 1] : 0;
    f[i] = x < lim ? L.F[x + 1] : 0;
    tot += v[i];
    ftot += f[i];
  }
  const u64 incl = wave_incl_sum(tot);
  const u32 fincl = wave_incl_sum(ftot);
  if (lane == 63) {
    S.wsum[wave] = incl;
    S.fsum[wave] = fincl;
  }
  __syncthreads();  // also: every thread has read its sizes before they are overwritten
  u64 run = incl - tot;
  u32 frun = fincl - ftot;
  for (u32 w = 0; w < wave; w++) {
    run += S.wsum[w];
    frun += S.fsum[w];
  }
  if (tid == 0) {
    L.P[0] = 0;
    L.F[0] = 0;
  }
  bool big = false;
#pragma unroll
  for (u32 i = 0; i < kStagePer; i++) {
    const u32 x = kStagePer * tid + i;
    run += v[i];
    frun += f[i];
    big |= run > 0xffffffffull;
    if (x < lim) {
      L.P[x + 1] = (u32)run;
      L.F[x + 1] = (u16)frun;
    }
  }
  if (big) cut_fall_back(st);
  __syncthreads();
  // block ends: starts interleaved over the threads (neighbouring lanes read
  // neighbouring LDS words), each search started at the tile's mean block
  const u64 bs = ep.block_size;
  const u32 R = ep.interval;
  const bool stream_end = T0 + lim == m;
  const u32 est = cuts::mean_block(L, lim, bs);
  bool unres = false;
#pragma unroll 1
  for (u32 i = 0; i < kStartsPer; i++) {
    const u32 x = i * kCutThreads + tid;
    if (x >= te) break;
    u64 bytes = 0;
    u32 e = cuts::block_end(L, x, lim, stream_end, est, bs, R, &bytes);
    if (e == kUnresolved) {
      unres = true;
      e = x + 1;  // keeps every walk moving; the job runs again another way
    }
    L.nx[x] = (u16)e;
    b.bnext[T0 + x] = e;
    b.bbytes[T0 + x] = (u32)bytes;
  }
  if (unres) cut_fall_back(st);
  __syncthreads();
  const u32 C = cuts::entry_points(L, lim, te, stream_end, bs);
  if (C == 0 || C > kCutCand) {
    if (tid == 0) cut_fall_back(st);
    return;
  }
  if (tid == 0) cb.cand[t] = C;
  for (u32 c = tid; c < C; c += kCutThreads) cb.ftab[t * kCutCand + c] = cuts::walk(L, c, te);
}

constexpr u32 kEmitThreads = 256;

// cut_emit.  Tile t = blockIdx.x: the real chain through the tile from its
// entry point (one thread, over the block ends cut_tiles left in bnext), then
// each block's descriptor.
__global__ __launch_bounds__(kEmitThreads) void k_cut_emit(Kept k, const u32* m_dev, EncBufs b, CutBufs cb,
                                                           u32 nbo_cap, DevStatus* st) {
  __shared__ u16 nx[kCutTile];
  __shared__ u16 starts[kCutTile];
  __shared__ u32 s_n;
  const u64 m = *m_dev;
  const u64 t = blockIdx.x;
  const u64 T0 = t * kCutTile;
  if (T0 >= m || st->err) return;  // the tiles and the chain ran before this launch: the status is settled
  const u32 tid = threadIdx.x;
  const u32 te = (u32)min<u64>(kCutTile, m - T0);
  for (u32 x = tid; x < te; x += kEmitThreads) nx[x] = (u16)b.bnext[T0 + x];
  __syncthreads();
  if (tid == 0) {
    u32 j = cb.entry[t], n = 0;
    while (j < te) {
      starts[n++] = (u16)j;
      j = nx[j];
    }
    s_n = n;
  }
  __syncthreads();
  const u32 base = cb.base[t];
  for (u32 i = tid; i < s_n; i += kEmitThreads) {
    const u32 x = starts[i];
    const u64 q = T0 + x;
    const u64 id = (u64)base + i;
    if (id < nbo_cap) {
      b.ob_first[id] = (u32)q;
      b.ob_n[id] = nx[x] - x;
      b.ob_bytes[id] = b.bbytes[q];
      b.ob_forced[id] = (q > 0 && k.forced[q]) ? 1u : 0u;
    }
  }
}

constexpr u32 kChainThreads = 1024;
constexpr u32 kChainWaves = kChainThreads / 64;
constexpr u32 kChainRows = 24576;  // function entries held in LDS
constexpr u32 kChainTiles = 4096;  // tiles whose row offsets fit in LDS
constexpr u32 kChainPer = kChainTiles / kChainThreads;

struct ChainLds {
  u32 rows[kChainRows];         // the functions of all tiles, packed
  u32 roff[kChainTiles + 1];    // where each tile's function starts in rows
  u16 segE[kChainWaves][kCutCand];  // a segment's function: where the chain leaves it
  u32 segN[kChainWaves][kCutCand];  // and the blocks it starts on the way
  u32 s_e[kChainWaves], s_b[kChainWaves];
  u32 s_w[kChainWaves];
  u32 total;
};

// Entry points of tile t and its function at entry e (bad when e is not one
// of them), from the packed rows in LDS or from global memory.  Two
// instances rather than a choice per access: no pointer may point at either.
template <bool LDS>
__device__ __forceinline__ u32 chain_points(const ChainLds& S, const CutBufs& cb, u32 t) {
  if constexpr (LDS) return S.roff[t + 1] - S.roff[t];
  else return cb.cand[t];
}
template <bool LDS>
__device__ __forceinline__ u32 chain_row(const ChainLds& S, const CutBufs& cb, u32 t, u32 e, bool& bad) {
  if (e >= chain_points<LDS>(S, cb, t)) {
    bad = true;
    return 0;
  }
  if constexpr (LDS) return S.rows[S.roff[t] + e];
  else return cb.ftab[(u64)t * kCutCand + e];
}

// Segments of Ls tiles, one per wave: the segment's function over the entry
// points of its first tile; then the segments in order (thread 0); then each
// segment's tiles from its real entry point.
template <bool LDS>
__device__ void chain_compose(ChainLds& S, const CutBufs& cb, u32 T, bool& bad) {
  const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const u32 Ls = (T + kChainWaves - 1) / kChainWaves;
  const u32 t0 = wave * Ls, t1 = min(T, t0 + Ls);
  if (t0 < t1) {
    const u32 C0 = wave == 0 ? 1 : chain_points<LDS>(S, cb, t0);
    for (u32 c = lane; c < C0; c += 64) {
      u32 e = c, n = 0;
      for (u32 t = t0; t < t1; t++) {
        const u32 v = chain_row<LDS>(S, cb, t, e, bad);
        n += v >> 12;
        e = v & 0xfff;
      }
      S.segE[wave][c] = (u16)e;
      S.segN[wave][c] = n;
    }
  }
  __syncthreads();
  if (tid == 0) {
    u32 e = 0, B = 0;
    bool whole = true;
    for (u32 s = 0; s < kChainWaves; s++) {
      const u32 u0 = s * Ls;
      if (u0 >= T) break;
      S.s_e[s] = e;
      S.s_b[s] = B;
      const u32 C0 = s == 0 ? 1 : chain_points<LDS>(S, cb, u0);
      if (e >= C0) {
        whole = false;
        break;
      }
      B += S.segN[s][e];
      e = S.segE[s][e];
    }
    if (!whole || e != 0) bad = true;  // the chain ends with the stream
    S.total = B;
  }
  __syncthreads();
  if (lane == 0 && t0 < t1) {
    u32 e = S.s_e[wave], B = S.s_b[wave];
    for (u32 t = t0; t < t1; t++) {
      cb.entry[t] = e;
      cb.base[t] = B;
      const u32 v = chain_row<LDS>(S, cb, t, e, bad);
      B += v >> 12;
      e = v & 0xfff;
    }
  }
}

// The functions of all tiles composed: each tile's entry point and the
// blocks before it, the block count (capacity-checked like a scan's total)
// and the sentinels after the last block.
__global__ __launch_bounds__(kChainThreads) void k_cut_chain(const u32* m_dev, CutBufs cb, EncBufs b, u32 nbo_cap,
                                                             DevStatus* st) {
  __shared__ ChainLds S;
  if (st->err) return;  // the tiles ran before this launch: the status is settled
  const u32 tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const u64 m = *m_dev;
  const u32 T = (u32)((m + kCutTile - 1) / kCutTile);
  if (T == 0) {
    if (tid == 0) {
      st->n_blocks_out = 0;
      b.ob_forced[0] = 0;
      b.ob_bytes[0] = 0;
    }
    return;
  }
  // the functions in LDS when they fit: row offsets by a scan of the counts
  bool lds = T <= kChainTiles;
  if (lds) {
    u32 c[kChainPer], sum = 0;
#pragma unroll
    for (u32 i = 0; i < kChainPer; i++) {
      const u32 t = kChainPer * tid + i;
      c[i] = t < T ? cb.cand[t] : 0;
      sum += c[i];
    }
    const u32 incl = wave_incl_sum(sum);
    if (lane == 63) S.s_w[wave] = incl;
    __syncthreads();
    u32 run = incl - sum;
    for (u32 w = 0; w < wave; w++) run += S.s_w[w];
#pragma unroll
    for (u32 i = 0; i < kChainPer; i++) {
      S.roff[kChainPer * tid + i] = run;
      run += c[i];
    }
    if (tid == kChainThreads - 1) S.roff[kChainTiles] = run;
    __syncthreads();
    lds = S.roff[kChainTiles] <= kChainRows;
    if (lds) {
      // every thread copies entries (t, c) = (g / 64, g % 64) for g = tid +
      // 1024 j, eight loads in flight; entry points past 64 after that
      const u32 G = 64 * T;
      for (u32 g0 = tid; g0 < G; g0 += 8 * kChainThreads) {
        u32 v[8];
#pragma unroll
        for (u32 j = 0; j < 8; j++) {
          const u32 g = g0 + j * kChainThreads, tt = g >> 6, c = g & 63;
          v[j] = g < G && c < S.roff[tt + 1] - S.roff[tt] ? cb.ftab[(u64)tt * kCutCand + c] : 0u;
        }
#pragma unroll
        for (u32 j = 0; j < 8; j++) {
          const u32 g = g0 + j * kChainThreads, tt = g >> 6, c = g & 63;
          if (g < G && c < S.roff[tt + 1] - S.roff[tt]) S.rows[S.roff[tt] + c] = v[j];
        }
      }
      for (u32 tt = wave; tt < T; tt += kChainWaves)
        for (u32 c = 64 + lane; c < S.roff[tt + 1] - S.roff[tt]; c += 64)
          S.rows[S.roff[tt] + c] = cb.ftab[(u64)tt * kCutCand + c];
    }
    __syncthreads();
  }
  bool bad = false;
  if (lds) chain_compose<true>(S, cb, T, bad);
  else chain_compose<false>(S, cb, T, bad);
  if (bad) set_err(st, ERR_UNSUPPORTED, SITE_SCAN);
  if (tid == 0) {
    u32 tot = S.total;
    if (tot > nbo_cap) {  // the count outgrew its capacity: report it, carry on with none
      st->need_blocks = tot;
      tot = 0;
      set_err(st, ERR_OVERFLOW, SITE_CAPACITY);
    }
    st->n_blocks_out = tot;
    b.ob_forced[tot] = 0;
    b.ob_bytes[tot] = 0;
  }
}

}  // namespace

void run_block_cuts(KTimer& T, Kept k, const u32* m_dev, u32 m_cap, EncParams ep, EncBufs b, CutBufs cb, u32 nbo_cap,
                    DevStatus* st) {
  const u32 tiles = cut_tiles_for(m_cap);
  KLAUNCH(T, "cut_tiles", k_cut_tiles, tiles, kCutThreads, 0, k, m_dev, ep, b, cb, st);
  KLAUNCH(T, "cut_chain", k_cut_chain, 1, kChainThreads, 0, m_dev, cb, b, nbo_cap, st);
  KLAUNCH(T, "cut_emit", k_cut_emit, tiles, kEmitThreads, 0, k, m_dev, b, cb, nbo_cap, st);
}

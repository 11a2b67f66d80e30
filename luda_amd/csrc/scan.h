// scan.h — the single-pass exclusive scan, shared by the units that fuse work
// into it: the item loads and a per-item epilogue are functors, so a small
// kernel that only produces a scan's input, or only scatters by its result,
// runs inside the scan instead of as a launch of its own.
//
// Decoupled look-back: workgroup t scans tile t (4096 items), publishes the
// tile's sum, then wave 0 folds the published sums of the tiles before it (256
// at a time, four per lane) until it meets a tile that has published its
// inclusive prefix, publishes its own, and the workgroup writes its items.  A
// kernel boundary costs several microseconds here, so one launch per scan
// beats three (reduce, scan of tile sums, down-sweep); the tiles are large and
// the look-back wide because every tile reaches its look-back at about the
// same time, and a round trip to the words is about a microsecond.
//
// Look-back words: one u64 per tile in a buffer that the context keeps for
// its scans only (zeroed when allocated), holding status (2 bits: 1 = tile
// sum, 2 = inclusive prefix), the launch's epoch (14 bits; a word of an
// earlier launch never matches) and the value (48 bits).  Every access is a
// relaxed atomic at agent scope: the value travels in the same word as its
// status, so no other ordering is needed, and the outputs are read only by
// later kernels.  Workgroup t waits only on tiles below t, dispatched before
// it; the wait is bounded all the same (an error, never a hang).
#pragma once

#include "engine.h"

// Internal linkage: each unit that includes this gets its own kernels.
namespace lbscan {
namespace {

constexpr u32 kThreads = 256;
constexpr u32 kWaves = kThreads / 64;
constexpr u32 kItems = 16;
constexpr u32 kTile = kThreads * kItems;  // item t * kTile + i * kThreads + thread: loads and stores coalesce
constexpr u32 kBack = 4;                  // look-back words per lane: 256 tiles per round trip
constexpr u64 kValMask = (1ull << 48) - 1;
constexpr u32 kStAgg = 1, kStPre = 2;
constexpr u32 kSpinMax = 1u << 22;

__device__ __forceinline__ u64 word(u32 status, u32 epoch, u64 v) {
  return ((u64)status << 62) | ((u64)(epoch & 0x3fff) << 48) | (v & kValMask);
}

// Wave 0 of tile t (t >= 1): the sum of all items before the tile.  Lane l
// reads the words of tiles j - 4l - k (k < 4); tiles below 0 count as an
// inclusive prefix of 0.  Scanning tiles downwards, a lane's part ends at its
// first inclusive prefix.  When every lane up to the first one holding an
// inclusive prefix is ready, their values end the walk; when no lane holds
// one and all are ready, the 256 tile sums are added and the walk moves on.
template <typename Tout>
__device__ Tout look_back(const ScanArgs& a, u64 t) {
  const u32 lane = threadIdx.x & 63;
  Tout ex = 0;
  long long j = (long long)t - 1;
  u32 spins = 0;
  for (;;) {
    u64 w[kBack];
#pragma unroll
    for (u32 k = 0; k < kBack; k++) {
      const long long p = j - (long long)(kBack * lane + k);
      w[k] = p >= 0 ? __hip_atomic_load(a.words + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                    : word(kStPre, a.epoch, 0);
    }
    bool seen = false, ready = true;
    Tout sum = 0;
#pragma unroll
    for (u32 k = 0; k < kBack; k++) {
      const u32 s = ((w[k] >> 48) & 0x3fff) == (a.epoch & 0x3fff) ? (u32)(w[k] >> 62) : 0u;
      if (!seen) {
        ready &= s != 0;
        sum += (Tout)(w[k] & kValMask);
        seen = s == kStPre;
      }
    }
    const u64 pre = __ballot(seen);
    const u64 rdy = __ballot(ready);
    if (pre) {
      const u32 f = (u32)__builtin_ctzll(pre);
      const u64 need = f == 63 ? ~0ull : ((2ull << f) - 1);
      if ((rdy & need) == need) return ex + wave_last(wave_incl_sum(lane <= f ? sum : Tout(0)));
    } else if (rdy == ~0ull) {
      ex += wave_last(wave_incl_sum(sum));
      j -= 64 * kBack;
      continue;
    }
    if (++spins > kSpinMax) {
      if (lane == 0) set_err(a.st, ERR_UNSUPPORTED, SITE_SCAN);
      return ex;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

// The tile a workgroup scans.  Workgroups of a launch are spread over the
// XCDs, which dispatch independently, so workgroup t - 1 may still wait for
// a CU behind another stream's kernels while t runs and waits on it; when
// the scans of several streams wait so on one another's unstarted tiles,
// none moves (worker slots running jobs side by side showed it: the look-
// backs gave up after their spin limit).  With a.ticket, a tile is taken by
// ticket first thing, so every tile below a workgroup's is held by one
// already running, which will publish; the workgroup that takes the
// launch's last ticket sets the counter back to 0 for the stream's next
// launch.  The ticket costs an atomic on one word per workgroup (about
// 15 us on a 256 MiB job's merge), so a context alone on the device keeps
// tile t = workgroup t.  Called by every thread.
__device__ __forceinline__ u32 take_ticket(const ScanArgs& a, u32* s_ticket) {
  if (!a.ticket) return blockIdx.x;
  if (threadIdx.x == 0) {
    const u32 t = __hip_atomic_fetch_add(a.ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t + 1 == gridDim.x) __hip_atomic_store(a.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *s_ticket = t;
  }
  __syncthreads();
  return *s_ticket;
}

// Publishes tile t's sum, then (t >= 1) looks back and publishes its
// inclusive prefix; returns the tile's exclusive prefix.  Wave 0 only.
// 64-bit sums must stay below 2^48 (byte offsets do); larger ones are errors.
template <typename Tout>
__device__ Tout publish(const ScanArgs& a, u64 t, Tout agg) {
  const u32 lane = threadIdx.x & 63;
  if (sizeof(Tout) == 8 && agg > kValMask) {
    if (lane == 0) set_err(a.st, ERR_UNSUPPORTED, SITE_SCAN);
    agg = (Tout)kValMask;
  }
  if (lane == 0)
    __hip_atomic_store(a.words + t, word(t ? kStAgg : kStPre, a.epoch, agg), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  if (t == 0) return 0;
  const Tout ex = look_back<Tout>(a, t);
  u64 inc = (u64)(Tout)(ex + agg);
  if (sizeof(Tout) == 8 && (ex > kValMask || inc > kValMask)) {
    if (lane == 0) set_err(a.st, ERR_UNSUPPORTED, SITE_SCAN);
    inc = kValMask;
  }
  if (lane == 0)
    __hip_atomic_store(a.words + t, word(kStPre, a.epoch, inc), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return ex;
}

// Items from an array.
template <typename Tin>
struct ArrLoad {
  const Tin* p;
  __device__ Tin operator()(u64 i) const { return p[i]; }
};

// No epilogue.  An epilogue gives, for every item i < n with value v and
// exclusive prefix ex, pre(i, v) (its loads: a thread issues those of all its
// items before any item's stores) and item(i, v, ex, pre); then done(total, n)
// once, after the guard, in the last tile.
struct NoEpi {
  struct Pre {};
  template <typename Tin>
  __device__ Pre pre(u64, Tin) const {
    return {};
  }
  template <typename Tin, typename Tout>
  __device__ void item(u64, Tin, Tout, const Pre&) const {}
  template <typename Tout>
  __device__ void done(Tout, u64) const {}
};

// out may be null (the epilogue is the only consumer).  Item x of tile t is
// row x / kThreads, thread x % kThreads: every row scans by waves (DPP), and
// the 128 wave sums of a tile, in item order, by wave 0.
template <typename Tin, typename Tout, typename Load, typename Epi>
__global__ __launch_bounds__(kThreads) void k_scan(Load ld, Tout* out, const u32* n_dev, u32 n_add, Tout* total,
                                                   ScanArgs a, Epi epi) {
  constexpr u32 kPer = kItems * kWaves / 64;  // wave sums per lane of wave 0
  static_assert(kItems * kWaves % 64 == 0 && kPer <= 4, "wave 0 scans the wave sums");
  __shared__ Tout wsum[kItems * kWaves];  // (row, wave) in item order, then exclusive
  __shared__ Tout s_pre, s_agg;
  __shared__ u32 s_ticket;
  const u64 t = take_ticket(a, &s_ticket);
  const u64 n = (n_dev ? *n_dev : 0) + n_add;
  const u64 tiles = (n + kTile - 1) / kTile;
  if (t >= (tiles ? tiles : 1)) return;
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const u64 base = t * kTile + threadIdx.x;
  Tin v[kItems];
#pragma unroll
  for (u32 i = 0; i < kItems; i++) {
    const u64 x = base + (u64)i * kThreads;
    v[i] = x < n ? ld(x) : Tin(0);
  }
  Tout inc[kItems];
#pragma unroll
  for (u32 i = 0; i < kItems; i++) {
    inc[i] = wave_incl_sum((Tout)v[i]);
    if (lane == 63) wsum[i * kWaves + wave] = inc[i];
  }
  __syncthreads();
  if (wave == 0) {
    Tout x[kPer], sx = 0;
#pragma unroll
    for (u32 k = 0; k < kPer; k++) {
      x[k] = wsum[kPer * lane + k];
      sx += x[k];
    }
    const Tout ip = wave_incl_sum(sx);
    Tout e = ip - sx;
#pragma unroll
    for (u32 k = 0; k < kPer; k++) {
      wsum[kPer * lane + k] = e;
      e += x[k];
    }
    const Tout agg = wave_last(ip);
    const Tout ex = publish<Tout>(a, t, agg);
    if (lane == 0) {
      s_pre = ex;
      s_agg = agg;
    }
  }
  __syncthreads();
  const Tout ex = s_pre;
  typename Epi::Pre pf[kItems];
#pragma unroll
  for (u32 i = 0; i < kItems; i++) {
    const u64 x = base + (u64)i * kThreads;
    if (x < n) pf[i] = epi.pre(x, v[i]);
  }
#pragma unroll
  for (u32 i = 0; i < kItems; i++) {
    const u64 x = base + (u64)i * kThreads;
    if (x < n) {
      const Tout pre = ex + wsum[i * kWaves + wave] + inc[i] - (Tout)v[i];
      if (out) out[x] = pre;
      epi.item(x, v[i], pre, pf[i]);
    }
  }
  if (threadIdx.x == 0 && t + 1 >= tiles) {
    Tout tot = ex + s_agg;
    if (a.need && tot > (Tout)a.cap) {  // the count outgrew its capacity: report it, carry on with none
      *a.need = (u32)tot;
      tot = 0;
      set_err(a.st, ERR_OVERFLOW, SITE_CAPACITY);
    }
    if (total) *total = tot;
    epi.done(tot, n);
  }
}

// One launch over n = (n_dev ? *n_dev : 0) + n_add <= n_cap items.
template <typename Tin, typename Tout, typename Load, typename Epi = NoEpi>
void run(KTimer& T, const char* name, ScanState& S, Load ld, Tout* out, const u32* n_dev, u32 n_add, u32 n_cap,
         Tout* total, DevStatus* st, const ScanGuard* g = nullptr, Epi epi = Epi()) {
  const u64 tiles = scan_tiles_for(n_cap);
  if (tiles > S.tiles) throw_scan_capacity(tiles, S.tiles);
  ScanArgs a;
  a.words = S.words;
  a.ticket = S.tickets ? S.ticket : nullptr;
  a.epoch = scan_next_epoch(S, T.s);
  a.st = st;
  a.need = g ? g->need : nullptr;
  a.cap = g ? g->cap : 0;
  KLAUNCH(T, name, (k_scan<Tin, Tout, Load, Epi>), (u32)tiles, kThreads, 0, ld, out, n_dev, n_add, total, a, epi);
}

}  // namespace
}  // namespace lbscan

// k_scan.hip — exclusive prefix sums whose length lives in device memory, so
// that a job runs with no host round trip between stages; plus fills.
//
// Three passes over tiles of 2048 items: per-tile sums, one workgroup scanning
// the tile sums, and per-tile scans offset by the tile prefix.  Every kernel
// reads n from device memory and exits early past it, so launch shapes depend
// only on the host's capacity bound.
#include "engine.h"

namespace {

constexpr u32 kThreads = 256;
constexpr u32 kItems = 8;
constexpr u32 kTile = kThreads * kItems;
constexpr u32 kTopThreads = 1024;

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  return wave_incl_sum(v);
}

// Exclusive scan across the workgroup; returns this thread's exclusive prefix
// and the workgroup total in *total.  `wsum` holds one T per wave.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* wsum, T* total) {
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  T inc = wave_incl_scan(v);
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  if (wave == 0) {
    T x = lane < nw ? wsum[lane] : T(0);
    T xi = wave_incl_scan(x);
    if (lane < nw) wsum[lane] = xi - x;
    if (lane == nw - 1) wsum[nw] = xi;
  }
  __syncthreads();
  T r = inc - v + wsum[wave];
  *total = wsum[nw];
  __syncthreads();
  return r;
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(kThreads) void k_scan_reduce(const Tin* in, const u32* n_dev, u32 n_add, Tout* tsum) {
  __shared__ Tout wsum[kThreads / 64 + 1];
  const u64 n = (n_dev ? *n_dev : 0) + n_add;
  const u64 tiles = (n + kTile - 1) / kTile;
  for (u64 t = blockIdx.x; t < tiles; t += gridDim.x) {
    Tout s = 0;
    const u64 base = t * kTile + (u64)threadIdx.x * kItems;
#pragma unroll
    for (u32 i = 0; i < kItems; i++)
      if (base + i < n) s += (Tout)in[base + i];
    Tout tot;
    block_excl_scan(s, wsum, &tot);
    if (threadIdx.x == 0) tsum[t] = tot;
  }
}

template <typename Tout>
__global__ __launch_bounds__(kTopThreads) void k_scan_tiles(Tout* tsum, const u32* n_dev, u32 n_add, Tout* total) {
  __shared__ Tout wsum[kTopThreads / 64 + 1];
  const u64 n = (n_dev ? *n_dev : 0) + n_add;
  const u64 tiles = (n + kTile - 1) / kTile;
  const u64 per = (tiles + kTopThreads - 1) / kTopThreads;
  const u64 b = (u64)threadIdx.x * per;
  Tout s = 0;
  for (u64 i = b; i < b + per && i < tiles; i++) s += tsum[i];
  Tout tot;
  Tout pre = block_excl_scan(s, wsum, &tot);
  for (u64 i = b; i < b + per && i < tiles; i++) {
    Tout v = tsum[i];
    tsum[i] = pre;
    pre += v;
  }
  if (threadIdx.x == 0 && total) *total = tot;
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(kThreads) void k_scan_down(const Tin* in, Tout* out, const u32* n_dev, u32 n_add,
                                                        const Tout* tpre) {
  __shared__ Tout wsum[kThreads / 64 + 1];
  const u64 n = (n_dev ? *n_dev : 0) + n_add;
  const u64 tiles = (n + kTile - 1) / kTile;
  for (u64 t = blockIdx.x; t < tiles; t += gridDim.x) {
    const u64 base = t * kTile + (u64)threadIdx.x * kItems;
    Tout v[kItems];
    Tout s = 0;
#pragma unroll
    for (u32 i = 0; i < kItems; i++) {
      v[i] = base + i < n ? (Tout)in[base + i] : Tout(0);
      s += v[i];
    }
    Tout tot;
    Tout pre = block_excl_scan(s, wsum, &tot) + tpre[t];
#pragma unroll
    for (u32 i = 0; i < kItems; i++) {
      if (base + i < n) out[base + i] = pre;
      pre += v[i];
    }
  }
}

// The whole scan in one workgroup of 1024 threads, chunk by chunk with a
// running carry: for short arrays one launch beats three.  In place is
// allowed (a chunk is read into registers before any of it is written).
constexpr u32 kSingleItems = 8;
constexpr u64 kSingleMax = 16384;
template <typename Tin, typename Tout>
__global__ __launch_bounds__(kTopThreads) void k_scan_single(const Tin* in, Tout* out, const u32* n_dev, u32 n_add,
                                                             Tout* total) {
  __shared__ Tout wsum[kTopThreads / 64 + 1];
  const u64 n = (n_dev ? *n_dev : 0) + n_add;
  Tout carry = 0;
  for (u64 c0 = 0; c0 < n; c0 += (u64)kTopThreads * kSingleItems) {
    const u64 base = c0 + (u64)threadIdx.x * kSingleItems;
    Tout v[kSingleItems];
    Tout s = 0;
#pragma unroll
    for (u32 i = 0; i < kSingleItems; i++) {
      v[i] = base + i < n ? (Tout)in[base + i] : Tout(0);
      s += v[i];
    }
    Tout tot;
    Tout pre = block_excl_scan(s, wsum, &tot) + carry;
#pragma unroll
    for (u32 i = 0; i < kSingleItems; i++) {
      if (base + i < n) out[base + i] = pre;
      pre += v[i];
    }
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <typename T>
__global__ void k_fill(T* p, T v, u64 n) {
  GRID_STRIDE(i, n) p[i] = v;
}
__global__ void k_iota(u32* p, u64 n) {
  GRID_STRIDE(i, n) p[i] = (u32)i;
}

u32 grid_for(u64 n, u32 per_block) {
  u64 g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return (u32)g;
}

template <typename Tin, typename Tout>
void scan_impl(KTimer& T, const Tin* in, Tout* out, const u32* n_dev, u32 n_add, u32 n_cap, Tout* total,
               void* tmp) {
  if (n_cap <= kSingleMax) {
    KLAUNCH(T, "scan_single", (k_scan_single<Tin, Tout>), 1, kTopThreads, 0, in, out, n_dev, n_add, total);
    return;
  }
  const u64 tiles = ((u64)n_cap + kTile - 1) / kTile;
  const u32 g = grid_for(tiles, 1);
  Tout* ts = reinterpret_cast<Tout*>(tmp);
  KLAUNCH(T, "scan_reduce", (k_scan_reduce<Tin, Tout>), g, kThreads, 0, in, n_dev, n_add, ts);
  KLAUNCH(T, "scan_tiles", (k_scan_tiles<Tout>), 1, kTopThreads, 0, ts, n_dev, n_add, total);
  KLAUNCH(T, "scan_down", (k_scan_down<Tin, Tout>), g, kThreads, 0, in, out, n_dev, n_add, (const Tout*)ts);
}

}  // namespace

size_t scan_tmp_bytes(u64 n_cap) { return (((n_cap + kTile - 1) / kTile) + 1) * sizeof(u64); }

void scan_u32(KTimer& T, const u32* in, u32* out, const u32* n_dev, u32 n_add, u32 n_cap, u32* total, void* tmp) {
  scan_impl<u32, u32>(T, in, out, n_dev, n_add, n_cap, total, tmp);
}
void scan_u64(KTimer& T, const u64* in, u64* out, const u32* n_dev, u32 n_add, u32 n_cap, u64* total, void* tmp) {
  scan_impl<u64, u64>(T, in, out, n_dev, n_add, n_cap, total, tmp);
}
void scan_u32_to_u64(KTimer& T, const u32* in, u64* out, const u32* n_dev, u32 n_add, u32 n_cap, u64* total,
                     void* tmp) {
  scan_impl<u32, u64>(T, in, out, n_dev, n_add, n_cap, total, tmp);
}
void fill_u32(KTimer& T, u32* p, u32 v, u64 n) {
  if (!n) return;
  KLAUNCH(T, "fill", k_fill<u32>, grid_for(n, 256 * 4), 256, 0, p, v, n);
}
void fill_u8(KTimer& T, u8* p, u8 v, u64 n) {
  if (!n) return;
  KLAUNCH(T, "fill", k_fill<u8>, grid_for(n, 256 * 4), 256, 0, p, v, n);
}
void iota_u32(KTimer& T, u32* p, u64 n) {
  if (!n) return;
  KLAUNCH(T, "iota", k_iota, grid_for(n, 256 * 4), 256, 0, p, n);
}

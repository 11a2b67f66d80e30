// k_scan.hip — the plain prefix sums (items from an array, no epilogue) of
// the single-pass scan in scan.h, the context's look-back words, and fills.
#include "scan.h"

namespace {

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

// Zeroes the look-back words with a kernel, not a runtime memset, so that a
// stream captured into a graph replays it as an ordinary kernel node.
hipError_t zero_words(u64* p, u64 n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill<u64>, dim3(grid_for(n, 256 * 4)), dim3(256), 0, s, p, (u64)0, n);
  return hipGetLastError();
}

}  // namespace

u64 scan_tiles_for(u64 n_cap) { return std::max<u64>(1, (n_cap + lbscan::kTile - 1) / lbscan::kTile); }

hipError_t scan_reserve(ScanState& S, u64 n_cap, hipStream_t s) { return scan_reserve_tiles(S, scan_tiles_for(n_cap), s); }

hipError_t scan_reserve_tiles(ScanState& S, u64 tiles, hipStream_t s) {
  if (tiles <= S.tiles) return hipSuccess;
  if (S.words) {
    const hipError_t e = hipStreamSynchronize(s);  // no scan of the stream still reads the old words
    if (e != hipSuccess) return e;
  }
  scan_release(S);
  const u64 want = std::max<u64>(tiles, 4096);
  hipError_t e = hipMalloc(&S.words, (want + 1) * sizeof(u64));  // + the ticket word
  if (e != hipSuccess) {
    S.words = nullptr;
    return e;
  }
  e = zero_words(S.words, want + 1, s);
  if (e != hipSuccess) return e;
  S.tiles = want;
  S.ticket = reinterpret_cast<u32*>(S.words + want);
  S.epoch = 0;
  return hipSuccess;
}

hipError_t scan_attach(ScanState& S, u64* words, u64 words_n, hipStream_t s) {
  S.words = words;
  S.tiles = words_n - 1;  // the last word is the ticket
  S.ticket = reinterpret_cast<u32*>(words + S.tiles);
  S.epoch = 0;
  return zero_words(words, words_n, s);
}

void scan_release(ScanState& S) {
  if (S.words) hipFree(S.words);
  S.words = nullptr;
  S.tiles = 0;
  S.ticket = nullptr;
  S.epoch = 0;
}

u32 scan_next_epoch(ScanState& S, hipStream_t s) {
  if (++S.epoch >= 0x4000) {
    const hipError_t e = zero_words(S.words, S.tiles + 1, s);
    if (e != hipSuccess) launch_refused("scan words reset", e);
    S.epoch = 1;
  }
  return S.epoch;
}

void scan_u32(KTimer& T, ScanState& S, const u32* in, u32* out, const u32* n_dev, u32 n_add, u32 n_cap, u32* total,
              DevStatus* st, const ScanGuard* g) {
  lbscan::run<u32, u32>(T, "scan", S, lbscan::ArrLoad<u32>{in}, out, n_dev, n_add, n_cap, total, st, g);
}
void scan_u64(KTimer& T, ScanState& S, const u64* in, u64* out, const u32* n_dev, u32 n_add, u32 n_cap, u64* total,
              DevStatus* st) {
  lbscan::run<u64, u64>(T, "scan", S, lbscan::ArrLoad<u64>{in}, out, n_dev, n_add, n_cap, total, st);
}
void scan_u32_to_u64(KTimer& T, ScanState& S, const u32* in, u64* out, const u32* n_dev, u32 n_add, u32 n_cap,
                     u64* total, DevStatus* st) {
  lbscan::run<u32, u64>(T, "scan", S, lbscan::ArrLoad<u32>{in}, out, n_dev, n_add, n_cap, total, st);
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

// stage_values.h — where the data block writer (k_write_blocks) puts the
// values of a block into its LDS stage.  Unaligned LDS stores cost several
// aligned ones, so:
//   - a value of at most 8 chunks of 16 bytes is stored by its entry's lane
//     as whole aligned dwords, realigned in registers; the dword slots run
//     from dst & ~3 to dst + vl rounded up, so up to 3 bytes before the
//     value and up to 3 after it get junk.  Those bytes are the entry's own
//     header and key (at least 4 bytes) and the next entry's header or the
//     restart array (at least 3), which the writer stores after the values
//     (one wave's LDS stores land in program order);
//   - longer values are spread over the wave by the 16-byte aligned slots of
//     the stage that they touch: a slot inside a value is one aligned 16-byte
//     store, the partial slots at a value's two ends are stored byte-exact.
// Written once for the device and for a host build (plain C++, bounds
// checked by the caller's accessors) that a CPU test runs against a
// restatement of BlockBuilder's layout.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define SV_FN __host__ __device__ __forceinline__
#else
#define SV_FN inline
#endif

namespace sv {

// ({hi, lo} >> 8 s)[31:0], s in 0..3: v_alignbyte_b32
SV_FN uint32_t alignbyte(uint32_t hi, uint32_t lo, uint32_t s) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbyte(hi, lo, s);
#else
  return (uint32_t)(((((uint64_t)hi) << 32) | lo) >> (8 * (s & 3)));
#endif
}

// Four little-endian words of a 16-byte chunk.
struct Chunk {
  uint32_t w[4];
};

// The value's vl <= 128 bytes, chunk t being bytes [16 t, 16 t + 16) of the
// source (read up to 15 bytes past the value), stored from stage byte
// address dst: put(k, word) stores dword slot k (byte address
// (dst & ~3) + 4 k).  The first KH chunks come in `first`; load(t) gives the
// later ones, asked for KH at a time.
template <uint32_t KH, typename Load, typename Put>
SV_FN void store_value(uint32_t dst, uint32_t vl, const Chunk (&first)[KH], Load load, Put put) {
  const uint32_t nch = (vl + 15) >> 4;
  const uint32_t a = dst & 3;
  const uint32_t nslot = (a + vl + 3) >> 2;
  // slot k = bytes 4 - a.. of source dword k - 1, then source dword k
  uint32_t prev = 0;
  auto run = [&](const Chunk(&c)[KH], uint32_t t0) {
#pragma unroll
    for (uint32_t t = 0; t < KH; t++) {
      if (t0 + t < nch) {
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) {
          const uint32_t k = 4 * (t0 + t) + j;
          const uint32_t w = c[t].w[j];
          if (k < nslot) put(k, a ? alignbyte(w, prev, 4 - a) : w);
          prev = w;
        }
      }
    }
  };
  run(first, 0);
#pragma unroll
  for (uint32_t t0 = KH; t0 < 8; t0 += KH) {
    if (t0 < nch) {
      Chunk late[KH];
#pragma unroll
      for (uint32_t t = 0; t < KH; t++)
        if (t0 + t < nch) late[t] = load(t0 + t);
      run(late, t0);
    }
  }
  if (4 * nch < nslot) put(4 * nch, alignbyte(0u, prev, 4 - a));  // a != 0 here
}

// Slots of a value at stage byte offset eo (stage byte x is 16-byte aligned
// when (g + x) % 16 == 0) of evl > 0 bytes.
SV_FN uint32_t spread_slots(uint32_t g, uint32_t eo, uint32_t evl) {
  return ((g + eo + evl + 15) >> 4) - ((g + eo) >> 4);
}

// Its k-th slot: the slot's first byte x and the value's bytes [lo, hi) in
// it (offsets in the stage); full when they are the whole slot.
struct Slot {
  int x, lo, hi;
  bool full;
};
SV_FN Slot spread_slot(uint32_t g, int eo, int evl, uint32_t k) {
  Slot s;
  s.x = eo - (int)((g + (uint32_t)eo) & 15) + 16 * (int)k;
  s.lo = s.x > eo ? s.x : eo;
  s.hi = s.x + 16 < eo + evl ? s.x + 16 : eo + evl;
  s.full = s.lo == s.x && s.hi == s.x + 16;
  return s;
}

}  // namespace sv

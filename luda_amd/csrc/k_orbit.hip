// k_orbit.hip — the nodes reached from node 0 by a successor function nxt
// (i < nxt[i] <= n, nxt[i] - i <= step).  This is how the encoder finds, in
// parallel, the exact greedy cut points that TableBuilder makes one entry at a
// time: where each data block starts (table/table_builder.cc:119-123) and
// where each output table starts (db/db_impl.cc:1173-1180).
//
// The nodes are cut into chunks of C.  The orbit enters chunk c at one of its
// first `step` nodes, so each chunk is summarised by a function from entry
// offset to the entry offset of chunk c+1 (computed by pointer jumping in
// LDS).  Those functions are composed in groups of G, level by level, until
// one is left; the entry offsets are then pushed back down the levels, and
// each chunk marks its own orbit nodes.
#include <string.h>

#include "engine.h"

namespace {

constexpr u16 kEnd = 0xffff;
constexpr u32 kMaxOrbitLevels = 16;
constexpr u32 kThreads = 256;

__device__ __forceinline__ u64 level_items(u64 n, u32 C, u32 G, u32 level) {
  u64 it = (n + C - 1) / C;
  for (u32 l = 0; l < level; l++) it = (it + G - 1) / G;
  return it;
}

// dst[i] = src[base + i] (fill past n) for i < cnt, by the whole workgroup,
// eight loads in flight per thread before any is stored.
__device__ __forceinline__ void stage_u32(u32* dst, const u32* src, u64 base, u32 cnt, u64 n, u32 fill) {
  constexpr u32 kBatch = 8;
  for (u32 i0 = 0; i0 < cnt; i0 += blockDim.x * kBatch) {
    u32 v[kBatch];
#pragma unroll
    for (u32 t = 0; t < kBatch; t++) {
      const u32 i = i0 + t * blockDim.x + threadIdx.x;
      v[t] = (i < cnt && base + i < n) ? src[base + i] : fill;
    }
#pragma unroll
    for (u32 t = 0; t < kBatch; t++) {
      const u32 i = i0 + t * blockDim.x + threadIdx.x;
      if (i < cnt) dst[i] = v[t];
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_orbit_chunk(const u32* nxt, const u32* n_dev, const u32* step_dev,
                                                           u32 C, u32 Wc, u16* f, DevStatus* st) {
  extern __shared__ u32 ex[];
  const u64 n = *n_dev;
  const u32 step = *step_dev;
  if (step > Wc) {
    if (blockIdx.x == 0 && threadIdx.x == 0) set_err(st, ERR_OVERFLOW, SITE_BLOCK_ENTRIES);
    return;
  }
  const u32 Wd = step ? step : 1;
  const u64 nch = (n + C - 1) / C;
  for (u64 c = blockIdx.x; c < nch; c += gridDim.x) {
    const u64 c0 = c * C;
    const u64 cend = c0 + C;
    stage_u32(ex, nxt, c0, C, n, (u32)n);
    __syncthreads();
    // pointer jumping until every pointer has left the chunk (at most
    // log2(C) + 1 rounds; fewer when the paths through the chunk are short)
    for (u32 r = 1; r <= C; r <<= 1) {
      int inside = 0;
      for (u32 i = threadIdx.x; i < C; i += blockDim.x) {
        const u32 v = ex[i];
        if (v < cend && v < n) {
          const u32 w = ex[v - c0];
          ex[i] = w;
          inside |= w < cend && w < n;
        }
      }
      if (!__syncthreads_or(inside)) break;
    }
    for (u32 x = threadIdx.x; x < Wd; x += blockDim.x) {
      u16 r = kEnd;
      if (c0 + x < n && x < C) {
        const u64 v = ex[x];
        if (v < n) {
          const u64 off = v - cend;
          if (off >= Wd) set_err(st, ERR_OVERFLOW, SITE_BLOCK_ENTRIES);
          else r = (u16)off;
        }
      }
      f[c * Wc + x] = r;
    }
    __syncthreads();
  }
}

// Level l -> l+1: compose G consecutive functions.
__global__ __launch_bounds__(kThreads) void k_orbit_compose(const u16* fin, u16* fout, const u32* n_dev,
                                                             const u32* step_dev, u32 C, u32 G, u32 Wc, u32 level) {
  extern __shared__ u16 tab[];
  const u64 n = *n_dev;
  const u32 step = *step_dev;
  if (step > Wc) return;  // k_orbit_chunk reported the overflow
  const u32 Wd = step ? step : 1;
  const u64 nin = level_items(n, C, G, level);
  const u64 nout = (nin + G - 1) / G;
  for (u64 o = blockIdx.x; o < nout; o += gridDim.x) {
    const u64 t0 = o * G;
    const u32 tc = (u32)min((u64)G, nin - t0);
    for (u32 i = threadIdx.x; i < tc * Wd; i += blockDim.x) {
      const u32 t = i / Wd, x = i % Wd;
      tab[i] = fin[(t0 + t) * Wc + x];
    }
    __syncthreads();
    for (u32 x = threadIdx.x; x < Wd; x += blockDim.x) {
      u32 y = x;
      for (u32 t = 0; t < tc && y != kEnd; t++) y = y < Wd ? tab[t * Wd + y] : kEnd;
      fout[o * Wc + x] = (u16)y;
    }
    __syncthreads();
  }
}

// Level l+1 entries -> level l entries.
__global__ __launch_bounds__(kThreads) void k_orbit_expand(const u16* fin, const u32* eout, u32* ein, const u32* n_dev,
                                                            const u32* step_dev, u32 C, u32 G, u32 Wc, u32 level) {
  extern __shared__ u16 tab[];
  const u64 n = *n_dev;
  const u32 step = *step_dev;
  if (step > Wc) return;  // k_orbit_chunk reported the overflow
  const u32 Wd = step ? step : 1;
  const u64 nin = level_items(n, C, G, level);
  const u64 nout = (nin + G - 1) / G;
  for (u64 o = blockIdx.x; o < nout; o += gridDim.x) {
    const u64 t0 = o * G;
    const u32 tc = (u32)min((u64)G, nin - t0);
    for (u32 i = threadIdx.x; i < tc * Wd; i += blockDim.x) {
      const u32 t = i / Wd, x = i % Wd;
      tab[i] = fin[(t0 + t) * Wc + x];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      u32 y = eout[o];
      for (u32 t = 0; t < tc; t++) {
        ein[t0 + t] = y;
        if (y != kEnd) y = y < Wd ? tab[t * Wd + y] : kEnd;
      }
    }
    __syncthreads();
  }
}

__global__ void k_orbit_root(u32* e) { *e = 0; }

// All levels above the chunks in one workgroup: compose level by level up to
// the single top function, then push entry offsets back down, with a
// workgroup barrier between levels (levels are small: the chunk functions
// number n / C).  Same tables and results as k_orbit_compose / _root /
// _expand.  With `staged`, every level's functions are kept in LDS (Wd
// apart) and the walks read them there; only the level-0 functions are read
// from global memory, once.
struct OrbitLevels {
  u32 levels;
  u64 level_offset[kMaxOrbitLevels];  // of each level's functions in ftab
  u64 entry_offset[kMaxOrbitLevels];  // of each level's entries in entries
};
__global__ __launch_bounds__(1024) void k_orbit_top(u16* ftab, u32* entries, const u32* n_dev, const u32* step_dev,
                                                    u32 C, u32 G, u32 Wc, OrbitLevels L, u32 lds_cap) {
  extern __shared__ u16 lt[];
  // u32 arithmetic throughout: every level's function count times Wd is
  // below n (C >= 2 Wd), and n < 2^32
  const u32 n = *n_dev;
  const u32 step = *step_dev;
  if (step > Wc) return;  // k_orbit_chunk reported the overflow
  const u32 Wd = step ? step : 1;
  const u32 tid = threadIdx.x;
  u32 items[kMaxOrbitLevels + 1], loff[kMaxOrbitLevels + 2];
  items[0] = n / C + (n % C != 0);
  loff[0] = 0;
#pragma unroll
  for (u32 l = 0; l <= kMaxOrbitLevels; l++) {
    if (l + 1 <= kMaxOrbitLevels) items[l + 1] = items[l] / G + (items[l] % G != 0);
    loff[l + 1] = loff[l] + (l <= L.levels ? items[l] * Wd : 0);
  }
  const bool staged = loff[L.levels + 1] <= lds_cap;
  if (staged) {
    // eight loads in flight per thread before any is stored
    const u32 t0 = items[0] * Wd;
    for (u32 i0 = 0; i0 < t0; i0 += 8 * blockDim.x) {
      u16 v[8];
#pragma unroll
      for (u32 t = 0; t < 8; t++) {
        const u32 i = i0 + t * blockDim.x + tid;
        const u32 o = i / Wd;
        v[t] = i < t0 ? ftab[(u64)o * Wc + (i - o * Wd)] : (u16)0;
      }
#pragma unroll
      for (u32 t = 0; t < 8; t++) {
        const u32 i = i0 + t * blockDim.x + tid;
        if (i < t0) lt[i] = v[t];
      }
    }
    __syncthreads();
  }
  for (u32 l = 0; l < L.levels; l++) {
    const u16* fin = ftab + L.level_offset[l];
    u16* fout = ftab + L.level_offset[l + 1];
    const u32 nin = items[l], nout = items[l + 1];
    const u32 lin = loff[l], lout = loff[l + 1];
    for (u32 i = tid; i < nout * Wd; i += blockDim.x) {
      const u32 o = i / Wd;
      const u32 x = i - o * Wd;
      const u32 t0 = o * G;
      const u32 tc = min(G, nin - t0);
      u32 y = x;
      if (staged) {
        for (u32 t = 0; t < tc && y != kEnd; t++) y = y < Wd ? lt[lin + (t0 + t) * Wd + y] : kEnd;
        lt[lout + i] = (u16)y;
      } else {
        for (u32 t = 0; t < tc && y != kEnd; t++) y = y < Wd ? fin[(u64)(t0 + t) * Wc + y] : kEnd;
      }
      fout[(u64)o * Wc + x] = (u16)y;
    }
    __syncthreads();
  }
  if (tid == 0) entries[L.entry_offset[L.levels]] = 0;
  __syncthreads();
  for (u32 l = L.levels; l-- > 0;) {
    const u16* fin = ftab + L.level_offset[l];
    const u32* eout = entries + L.entry_offset[l + 1];
    u32* ein = entries + L.entry_offset[l];
    const u32 nin = items[l], nout = items[l + 1];
    const u32 lin = loff[l];
    for (u32 o = tid; o < nout; o += blockDim.x) {
      const u32 t0 = o * G;
      const u32 tc = min(G, nin - t0);
      u32 y = eout[o];
      for (u32 t = 0; t < tc; t++) {
        ein[t0 + t] = y;
        if (y != kEnd) {
          if (staged) y = y < Wd ? lt[lin + (t0 + t) * Wd + y] : kEnd;
          else y = y < Wd ? fin[(u64)(t0 + t) * Wc + y] : kEnd;
        }
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void k_orbit_mark(const u32* nxt, const u32* n_dev, const u32* step_dev,
                                                          u32 C, u32 Wc, const u32* e0, u32* flag) {
  extern __shared__ u32 loc[];
  u8* fl = reinterpret_cast<u8*>(loc + C);
  const u64 n = *n_dev;
  if (*step_dev > Wc) return;  // k_orbit_chunk reported the overflow
  const u64 nch = (n + C - 1) / C;
  for (u64 c = blockIdx.x; c < nch; c += gridDim.x) {
    const u64 c0 = c * C;
    const u64 c1 = min(c0 + C, n);
    stage_u32(loc, nxt, c0, C, n, (u32)n);
    for (u32 i = threadIdx.x; i < C; i += blockDim.x) fl[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
      const u32 y = e0[c];
      if (y < C) {
        u64 node = c0 + y;
        while (node < c1) {
          fl[node - c0] = 1;
          node = loc[node - c0];
        }
      }
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < c1 - c0; i += blockDim.x) flag[c0 + i] = fl[i];
    __syncthreads();
  }
}

u32 grid_for(u64 n) {
  if (n < 1) n = 1;
  if (n > 8192) n = 8192;
  return (u32)n;
}

}  // namespace

OrbitPlan plan_orbit(u64 n_cap, u32 C, u32 Wc, u32 G) {
  OrbitPlan p;
  p.C = C;
  p.Wc = Wc;
  p.G = G;
  p.n_cap = n_cap;
  u64 items = (n_cap + C - 1) / C;
  if (items < 1) items = 1;
  u64 toff = 0;
  p.level_items.push_back(items);
  p.level_offset.push_back(0);
  toff += items * Wc;
  while (items > 1) {
    items = (items + p.G - 1) / p.G;
    p.level_items.push_back(items);
    p.level_offset.push_back(toff);
    toff += items * Wc;
  }
  p.levels = (u32)p.level_items.size() - 1;
  p.table_entries = toff;
  u64 e = 0;
  for (auto it : p.level_items) e += it;
  p.entry_entries = e + 1;
  return p;
}

void run_orbit(KTimer& T, const OrbitPlan& p, const u32* nxt, const u32* n_dev, const u32* step_dev, u32* flag,
               u16* ftab, u32* entries, DevStatus* st) {
  const u32 L = p.levels;
  std::vector<u64> eoff(L + 1);
  u64 acc = 0;
  for (u32 l = 0; l <= L; l++) {
    eoff[l] = acc;
    acc += p.level_items[l];
  }
  KLAUNCH(T, "orbit_chunk", k_orbit_chunk, grid_for(p.level_items[0]), kThreads, p.C * sizeof(u32), nxt, n_dev,
          step_dev, p.C, p.Wc, ftab, st);
  if (L + 1 <= kMaxOrbitLevels && p.level_items[0] <= 4096) {
    // few chunks: every level above them in one launch
    OrbitLevels lv;
    memset(&lv, 0, sizeof(lv));
    lv.levels = L;
    for (u32 l = 0; l <= L; l++) {
      lv.level_offset[l] = p.level_offset[l];
      lv.entry_offset[l] = eoff[l];
    }
    // every level's functions in LDS when the step found on the device allows
    const u32 lds_cap = 30 * 1024;  // u16 entries: 60 KiB
    KLAUNCH(T, "orbit_top", k_orbit_top, 1, 1024, lds_cap * sizeof(u16), ftab, entries, n_dev, step_dev, p.C, p.G,
            p.Wc, lv, lds_cap);
    KLAUNCH(T, "orbit_mark", k_orbit_mark, grid_for(p.level_items[0]), kThreads, p.C * sizeof(u32) + p.C, nxt, n_dev,
            step_dev, p.C, p.Wc, entries, flag);
    return;
  }
  for (u32 l = 0; l < L; l++)
    KLAUNCH(T, "orbit_compose", k_orbit_compose, grid_for(p.level_items[l + 1]), kThreads,
            (size_t)p.G * p.Wc * sizeof(u16), ftab + p.level_offset[l], ftab + p.level_offset[l + 1], n_dev, step_dev,
            p.C, p.G, p.Wc, l);
  KLAUNCH(T, "orbit_root", k_orbit_root, 1, 1, 0, entries + eoff[L]);
  for (u32 l = L; l-- > 0;)
    KLAUNCH(T, "orbit_expand", k_orbit_expand, grid_for(p.level_items[l + 1]), kThreads,
            (size_t)p.G * p.Wc * sizeof(u16), ftab + p.level_offset[l], entries + eoff[l + 1], entries + eoff[l],
            n_dev, step_dev, p.C, p.G, p.Wc, l);
  KLAUNCH(T, "orbit_mark", k_orbit_mark, grid_for(p.level_items[0]), kThreads, p.C * sizeof(u32) + p.C, nxt, n_dev,
          step_dev, p.C, p.Wc, entries, flag);
}

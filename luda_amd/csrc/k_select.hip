// k_select.hip — which merged entries survive, and where ShouldStopBefore
// starts a new output table.
//
// The drop rule of DBImpl::DoCompactionWork (db/db_impl.cc:1113-1147) only
// looks at an entry and its predecessor: last_sequence_for_key is the
// predecessor's sequence when the predecessor parsed and has the same user
// key, else kMaxSequenceNumber.  So
//   drop(p) = parsed(p) && ((parsed(p-1) && ukey(p-1) == ukey(p) && seq(p-1) <= S)
//                           || (type(p) == deletion && seq(p) <= S && IsBaseLevelForKey(ukey(p))))
// is evaluated for all p at once.  IsBaseLevelForKey (db/version_set.cc:1535-1554)
// walks monotone per-level pointers over sorted files; for non-decreasing keys
// that equals a binary search for the first file whose largest key is >= the
// key.  ShouldStopBefore (db/version_set.cc:1556-1579) only changes state when
// the key passes a grandparent's largest key; those crossings are found in
// parallel and the short sequential budget walk runs over them alone.
#include "engine.h"
#include "scan.h"
#include "select.h"

namespace {

__global__ void k_drop(const u32* perm, u32 n, Planes K, DropParams dp, u32* keep, u32* gcnt, u32* slot) {
  if (dp.ngp) GRID_STRIDE(g, (u64)dp.ngp + 1) slot[g] = kNoEvent;  // filled by the keep scan
  GRID_STRIDE(p, n) {
    const u32 e = perm[p];
    const u64 tag = K.tag[e];
    const u32 type = (u32)(tag & 0xff);
    const u64 seq = tag >> 8;
    bool drop = false;
    if (type <= 1) {
      if (p > 0) {
        const u32 ep = perm[p - 1];
        const u64 tp = K.tag[ep];
        if ((tp & 0xff) <= 1 && (tp >> 8) <= dp.S && ukey_eq(K, e, ep)) drop = true;
      }
      if (!drop && type == 0 && seq <= dp.S && is_base_level_for_key(K, e, dp)) drop = true;
    }
    keep[p] = drop ? 0u : 1u;
    if (dp.ngp) gcnt[p] = grandparents_below(K, e, dp);
  }
}

__global__ void k_gather(const u32* perm, u32 n, const u32* keep, const u32* kpos, Planes K, const u64* vptr,
                         const u32* vlen, Kept out) {
  GRID_STRIDE(p, n) {
    if (!keep[p]) continue;
    const u32 q = kpos[p];
    const u32 e = perm[p];
    for (u32 w = 0; w < K.W; w++) out.K.P[w * out.K.stride + q] = K.P[w * K.stride + e];
    out.K.klen[q] = K.klen[e];
    out.K.tag[q] = K.tag[e];
    out.vptr[q] = vptr[e];
    out.vlen[q] = vlen[e];
    out.forced[q] = 0;
  }
}

// The keep scan's epilogue: the positions p >= 1 where the grandparent count
// grows.  gcnt only grows along the merged order, so each such p has a count
// of its own and goes to slot[gcnt[p]]: the slots list the crossings in order.
struct StopEventsEpi {
  const u32* gcnt;
  u32* slot;
  struct Pre {
    bool cross;
  };
  __device__ Pre pre(u64 p, u32) const { return {slot && p > 0 && gcnt[p] > gcnt[p - 1]}; }
  __device__ void item(u64 p, u32, u32, const Pre& c) const {
    if (c.cross) slot[gcnt[p]] = (u32)p;
  }
  __device__ void done(u32, u64) const {}
};

constexpr u32 kStopThreads = 1024;
constexpr u32 kStopChunk = 2048;  // crossing slots per pass through LDS
constexpr u32 kNoCut = 0xffffffffu, kEmpty = 0xfffffffeu;

// One workgroup walks the crossing slots in order, kStopChunk at a time: all
// threads gather what each crossing adds to the overlap budget and where it
// would cut, thread 0 walks them as ShouldStopBefore would, and the cuts
// that fire mark the kept entry they land on.
__global__ __launch_bounds__(kStopThreads) void k_stop_apply(const u32* gcnt, const u32* kpos, u32 n,
                                                              DropParams dp, const u32* slot, Kept out,
                                                              DevStatus* st) {
  __shared__ u64 add[kStopChunk];
  __shared__ u32 cut[kStopChunk];
  const u32 nkept = st->n_kept;
  // Compaction::ShouldStopBefore: overlapped_bytes_ grows by every
  // grandparent passed after the first key and resets when a cut fires.
  u64 over = 0;
  u32 nstops = 0;
  for (u32 g0 = 1; g0 <= dp.ngp; g0 += kStopChunk) {
    const u32 cnt = min(kStopChunk, dp.ngp + 1 - g0);
    for (u32 i = threadIdx.x; i < cnt; i += blockDim.x) {
      const u32 p = slot[g0 + i];
      if (p == kNoEvent || p == 0 || p >= n) {
        add[i] = 0;
        cut[i] = kEmpty;
      } else {
        add[i] = dp.gp_prefix[gcnt[p]] - dp.gp_prefix[gcnt[p - 1]];
        cut[i] = kpos[p];
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (u32 i = 0; i < cnt; i++) {
        if (cut[i] == kEmpty) continue;
        over += add[i];
        if (over > dp.limit) {
          over = 0;
          nstops++;
        } else {
          cut[i] = kNoCut;
        }
      }
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < cnt; i += blockDim.x) {
      const u32 q = cut[i];
      if (q != kNoCut && q != kEmpty && q > 0 && q < nkept) out.forced[q] = 1;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) st->n_stops = nstops;
}

u32 grid_for(u64 n, u32 per_block) {
  u64 g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return (u32)g;
}

}  // namespace

void launch_drop(KTimer& T, const u32* perm, u32 n, Planes K, const DropParams& dp, u32* keep, u32* gcnt, u32* events,
                 DevStatus* st) {
  (void)st;
  if (!n) return;
  KLAUNCH(T, "drop", k_drop, grid_for(n, 256), 256, 0, perm, n, K, dp, keep, gcnt, events);
}

void scan_keep(KTimer& T, ScanState& S, const u32* keep, u32* kpos, u32 n, const DropParams& dp, const u32* gcnt,
               u32* events, u32 ev_cap, DevStatus* st) {
  (void)ev_cap;
  const bool stops = n && dp.ngp;
  lbscan::run<u32, u32>(T, "scan_keep", S, lbscan::ArrLoad<u32>{keep}, kpos, nullptr, n, n, &st->n_kept, st, nullptr,
                        StopEventsEpi{gcnt, stops ? events : nullptr});
}

void launch_gather(KTimer& T, const u32* perm, u32 n, const u32* keep, const u32* kpos, Planes K, const u64* vptr,
                   const u32* vlen, Kept out) {
  if (!n) return;
  KLAUNCH(T, "gather", k_gather, grid_for(n, 256), 256, 0, perm, n, keep, kpos, K, vptr, vlen, out);
}

void launch_stops(KTimer& T, const u32* gcnt, const u32* kpos, const u32* keep, u32 n, const DropParams& dp,
                  u32* events, u32 ev_cap, Kept out, DevStatus* st) {
  (void)keep;
  (void)ev_cap;
  if (!n || !dp.ngp) return;
  // events[g], g = 0..ngp: the crossing slots (cleared by the drop kernel, filled by the keep scan).
  KLAUNCH(T, "stop_apply", k_stop_apply, 1, kStopThreads, 0, gcnt, kpos, n, dp, (const u32*)events, out, st);
}

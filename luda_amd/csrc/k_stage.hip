// k_stage.hip — small host<->device transfers of a job (file descriptors,
// the host-built tables of phase B, the status word, table metadata and
// boundary keys) done by one kernel each, reading or writing the context's
// pinned host buffers directly.  A job needs four of them; as runtime copies
// they were seven, each with its own enqueue cost on the host.
#include "engine.h"

namespace {

constexpr u32 kStageThreads = 256;

// Copies n bytes; whole 16-byte pieces when src and dst are both 16-byte
// aligned, single bytes otherwise and for the tail.
__device__ __forceinline__ void copy_bytes(const u8* __restrict__ src, u8* __restrict__ dst, u64 n, u64 tid,
                                           u64 nth) {
  u64 n16 = 0;
  if ((((u64)src | (u64)dst) & 15) == 0) {
    n16 = n >> 4;
    const uint4* s4 = reinterpret_cast<const uint4*>(src);
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    for (u64 i = tid; i < n16; i += nth) d4[i] = s4[i];
  }
  for (u64 i = 16 * n16 + tid; i < n; i += nth) dst[i] = src[i];
}

__global__ __launch_bounds__(kStageThreads) void k_stage_in(const u8* __restrict__ src, u8* __restrict__ dst,
                                                            u64 bytes, const DevStatus* __restrict__ init,
                                                            DevStatus* st, StageFill fill) {
  const u64 tid = (u64)blockIdx.x * blockDim.x + threadIdx.x, nth = (u64)gridDim.x * blockDim.x;
  copy_bytes(src, dst, bytes, tid, nth);
  for (u64 i = tid; i < fill.n; i += nth) fill.p[i] = fill.v;
  if (init && blockIdx.x == 0) {
    const u32* a = reinterpret_cast<const u32*>(init);
    u32* b = reinterpret_cast<u32*>(st);
    for (u32 i = threadIdx.x; i < sizeof(DevStatus) / 4; i += blockDim.x) b[i] = a[i];
  }
}

// With done.flag set, the last workgroup to finish its copies stores
// done.seq into the flag (host memory) after every staged byte is visible
// at system scope, so the host can see the copies complete by reading one
// word, without asking the runtime.  done.ctr counts finished workgroups and
// is reset by the last one (the next staging kernel of the stream starts
// after this one ends).
__global__ __launch_bounds__(kStageThreads) void k_stage_out(u8* __restrict__ dst, StageSeg s0, StageSeg s1,
                                                             StageSeg s2, StageDone done) {
  const u64 tid = (u64)blockIdx.x * blockDim.x + threadIdx.x, nth = (u64)gridDim.x * blockDim.x;
  if (s0.bytes) copy_bytes(reinterpret_cast<const u8*>(s0.src), dst + s0.off, s0.bytes, tid, nth);
  if (s1.bytes) copy_bytes(reinterpret_cast<const u8*>(s1.src), dst + s1.off, s1.bytes, tid, nth);
  if (s2.bytes) copy_bytes(reinterpret_cast<const u8*>(s2.src), dst + s2.off, s2.bytes, tid, nth);
  if (done.flag) {
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
      const u32 prev = __hip_atomic_fetch_add(done.ctr, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (prev == gridDim.x - 1) {
        __hip_atomic_store(done.ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence_system();
        __hip_atomic_store(done.flag, done.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
  }
}

u32 stage_grid(u64 bytes) {
  const u64 g = (bytes / 16 + kStageThreads - 1) / kStageThreads;
  return (u32)std::min<u64>(std::max<u64>(g, 1), 256);
}

}  // namespace

static_assert(sizeof(DevStatus) % 4 == 0, "the status is copied as 32-bit words");

void launch_stage_in(KTimer& T, const u8* src, u8* dst, u64 bytes, const DevStatus* init, DevStatus* st,
                     StageFill fill) {
  KLAUNCH(T, "stage_in", k_stage_in, stage_grid(std::max<u64>(bytes, 4 * fill.n)), kStageThreads, 0, src, dst, bytes,
          init, st, fill);
}

void launch_stage_out(KTimer& T, u8* dst, StageSeg s0, StageSeg s1, StageSeg s2, StageDone done) {
  KLAUNCH(T, "stage_out", k_stage_out, stage_grid(std::max(s0.bytes, std::max(s1.bytes, s2.bytes))),
          kStageThreads, 0, dst, s0, s1, s2, done);
}

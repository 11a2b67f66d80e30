// engine.hip — host driver of the compaction engine and the C-ABI of
// include/luda.h.
//
// One call runs two device phases on the context's stream:
//   A  footers, index blocks and restart arrays of the inputs: how many data
//      blocks and entries there are (one small read-back);
//   B  decode, merge, drop, table cut and encode, sized from A, then one
//      read-back of the status and the output metadata.
// Everything in between is driven by device-resident counts.  A capacity
// guess that turns out too small (key width, output bytes, tables, filter
// windows) is reported by the kernels and the call re-runs phase B with the
// measured sizes.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <memory>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"
#include "luda.h"

namespace {

constexpr u64 kAlign = 256;
// user keys up to 2 KiB: the wide-key decoder holds one key per lane in LDS
// (64 x (8 * 256 + 8) bytes beside its block stage, within a CU's 160 KiB)
constexpr u32 kMaxKeyWords = 256;
constexpr u64 kMaxBlockSize = 8192;   // block_size handled on the device
constexpr u32 kMaxChunk = 8192;       // orbit chunk: its LDS footprint is 5 bytes per node
constexpr u32 kMaxAttempts = 5;

u64 align_up(u64 x, u64 a) { return (x + a - 1) / a * a; }
constexpr u64 align_up_c(u64 x, u64 a) { return (x + a - 1) / a * a; }

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Bump allocator over one device allocation; a dry run (base == nullptr)
// measures the bytes a plan needs.
struct Arena {
  u8* base = nullptr;
  u64 cap = 0;
  u64 used = 0;
  bool dry = false;
  template <typename T>
  T* take(u64 count) {
    const u64 off = align_up(used, kAlign);
    used = off + align_up(count * sizeof(T), kAlign);
    if (dry) return nullptr;
    return reinterpret_cast<T*>(base + off);
  }
};

// Device memory of a context family (a context and its worker slots):
// chunks from hipMalloc, carved into blocks by best fit, free neighbours
// merged again.  Jobs take their working buffers and output tables here
// rather than from hipMalloc and hipFree: a hipFree waits for the whole
// device, so one worker slot growing a buffer stalled every other slot, and
// each fresh output allocation cost 0.1-0.5 ms of host time on the job's
// path.  A chunk goes back to the runtime only when it is wholly free and
// the family would still hold kRetain bytes beyond what is in use.
class DevHeap {
 public:
  static constexpr u64 kAlign = 4096;
  static constexpr u64 kChunkMin = 1ull << 30;
  static constexpr u64 kRetain = 8ull << 30;

  ~DevHeap() {
    for (const Chunk& ch : chunks_)
      if (ch.base) hipFree(ch.base);
  }
  // nullptr when the device has no room; the calling thread's current
  // device is the family's
  void* take(u64 bytes) {
    const u64 need = align_up(std::max<u64>(bytes, 1), kAlign);
    std::lock_guard<std::mutex> lk(mu_);
    auto it = free_.lower_bound(need);
    if (it == free_.end()) {
      u64 size = std::max<u64>(align_up(need, 2ull << 20), kChunkMin);
      u8* base = nullptr;
      if (hipMalloc(&base, size) != hipSuccess) {
        (void)hipGetLastError();
        base = nullptr;
        size = align_up(need, 2ull << 20);
        if (hipMalloc(&base, size) != hipSuccess) {
          (void)hipGetLastError();
          return nullptr;
        }
      }
      chunks_.push_back(Chunk{base, size});
      blocks_[base] = Block{size, (u32)(chunks_.size() - 1), true};
      it = free_.emplace(size, base);
      held_ += size;
    }
    u8* p = it->second;
    const u64 have = it->first;
    free_.erase(it);
    Block& b = blocks_.at(p);
    b.free = false;
    if (have - need >= kAlign) {
      b.size = need;
      blocks_[p + need] = Block{have - need, b.chunk, true};
      free_.emplace(have - need, p + need);
    }
    in_use_ += b.size;
    return p;
  }
  // false: p is not a block of this heap in use
  bool give(void* q) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = blocks_.find(static_cast<u8*>(q));
    if (it == blocks_.end() || it->second.free) return false;
    in_use_ -= it->second.size;
    it->second.free = true;
    auto nx = std::next(it);
    if (nx != blocks_.end() && nx->second.free && nx->second.chunk == it->second.chunk) {
      erase_free(nx->first, nx->second.size);
      it->second.size += nx->second.size;
      blocks_.erase(nx);
    }
    if (it != blocks_.begin()) {
      auto pv = std::prev(it);
      if (pv->second.free && pv->second.chunk == it->second.chunk) {
        erase_free(pv->first, pv->second.size);
        pv->second.size += it->second.size;
        blocks_.erase(it);
        it = pv;
      }
    }
    Chunk& ch = chunks_[it->second.chunk];
    if (it->first == ch.base && it->second.size == ch.size && held_ - ch.size >= in_use_ + kRetain) {
      held_ -= ch.size;
      hipFree(ch.base);
      ch.base = nullptr;
      blocks_.erase(it);
      return true;
    }
    free_.emplace(it->second.size, it->first);
    return true;
  }
  u64 held() {
    std::lock_guard<std::mutex> lk(mu_);
    return held_;
  }

 private:
  struct Chunk {
    u8* base;
    u64 size;
  };
  struct Block {
    u64 size;
    u32 chunk;
    bool free;
  };
  void erase_free(u8* p, u64 size) {
    auto r = free_.equal_range(size);
    for (auto i = r.first; i != r.second; ++i)
      if (i->second == p) {
        free_.erase(i);
        return;
      }
  }
  std::mutex mu_;
  std::vector<Chunk> chunks_;
  std::map<u8*, Block> blocks_;     // every block of every chunk, by address
  std::multimap<u64, u8*> free_;    // free blocks by size
  u64 in_use_ = 0, held_ = 0;
};

// A context's working buffer, grown (with a quarter of headroom) from the
// family's heap.  Growing waits for the context's own stream first, so no
// queued kernel still uses the old block when another slot takes it.
struct DevBuf {
  std::shared_ptr<DevHeap> heap;
  hipStream_t s = nullptr;
  u8* p = nullptr;
  u64 cap = 0;
  hipError_t reserve(u64 need) {
    if (need <= cap) return hipSuccess;
    if (p) {
      const hipError_t e = hipStreamSynchronize(s);
      if (e != hipSuccess) return e;
      heap->give(p);
    }
    p = nullptr;
    cap = 0;
    u64 c = std::max<u64>(need + need / 4, 1 << 20);
    void* q = heap->take(c);
    if (!q) q = heap->take(c = need);
    if (!q) return hipErrorOutOfMemory;
    p = static_cast<u8*>(q);
    cap = c;
    return hipSuccess;
  }
  void release() {
    if (p) heap->give(p);
    p = nullptr;
    cap = 0;
  }
};

// Pinned host memory; d is the address kernels use for it.  Grows with half
// again as much headroom: hipHostFree waits for the whole device, which in
// the batched form stalls every worker slot.
struct PinnedBuf {
  u8* p = nullptr;
  u8* d = nullptr;
  u64 cap = 0;
  hipError_t reserve(u64 need) {
    if (need <= cap) return hipSuccess;
    const bool grow = p != nullptr;
    release();
    u64 c = std::max<u64>(grow ? need + need / 2 : need, 1 << 16);
    hipError_t e = hipHostMalloc(&p, c, hipHostMallocDefault);
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    void* dp = nullptr;
    e = hipHostGetDevicePointer(&dp, p, 0);
    if (e != hipSuccess || !dp) {
      hipHostFree(p);
      p = nullptr;
      return e != hipSuccess ? e : hipErrorInvalidValue;
    }
    d = static_cast<u8*>(dp);
    cap = c;
    return hipSuccess;
  }
  void release() {
    if (p) hipHostFree(p);
    p = nullptr;
    d = nullptr;
    cap = 0;
  }
};

const char* site_name(u32 s) {
  switch (s) {
    case SITE_FOOTER: return "table footer";
    case SITE_INDEX: return "index block";
    case SITE_BLOCK_TRAILER: return "data block restart array";
    case SITE_ENTRY: return "block entry";
    case SITE_GROUP_COUNT: return "restart group entry count";
    case SITE_SHORT_KEY: return "internal key shorter than 8 bytes";
    case SITE_KEY_WIDTH: return "key width";
    case SITE_UNSORTED: return "input not strictly sorted";
    case SITE_CAPACITY: return "capacity";
    case SITE_COMPRESSED: return "compressed block";
    case SITE_BLOCK_ENTRIES: return "block entries";
    case SITE_FILE_BLOCKS: return "table blocks";
    case SITE_SCAN: return "prefix sum";
    case SITE_CHECKSUM: return "block checksum";
    default: return "unknown";
  }
}

// Key -> its first 8W bytes as big-endian words.
void to_words(const u8* k, u64 len, u32 W, u64* out) {
  for (u32 w = 0; w < W; w++) {
    u64 v = 0;
    for (u32 t = 0; t < 8; t++) {
      const u64 pos = 8ull * w + t;
      v = (v << 8) | (pos < len ? k[pos] : 0u);
    }
    out[w] = v;
  }
}

int bytewise_cmp(const u8* a, u64 an, const u8* b, u64 bn) {
  const u64 m = std::min(an, bn);
  int r = m ? memcmp(a, b, m) : 0;
  if (r == 0) r = an < bn ? -1 : (an > bn ? 1 : 0);
  return r < 0 ? -1 : (r > 0 ? 1 : 0);
}
int ikey_cmp_host(const u8* a, u64 an, const u8* b, u64 bn) {
  int r = bytewise_cmp(a, an - 8, b, bn - 8);
  if (r) return r;
  u64 ta = 0, tb = 0;
  memcpy(&ta, a + an - 8, 8);
  memcpy(&tb, b + bn - 8, 8);
  return ta > tb ? -1 : (ta < tb ? 1 : 0);
}

u64 next_pow2(u64 x) {
  u64 p = 1;
  while (p < x) p <<= 1;
  return p;
}

// The stream the encoder reads, plus the counts it depends on.
struct EncodeSetup {
  EncParams ep;
  u32 m_cap;      // kept entries
  u32 nbo_cap;    // output blocks
  u32 nf_cap;     // output tables
  u32 nw_cap;     // filter windows
  u64 out_cap;    // output bytes
  bool size_cuts; // max_file_size can cut a table before the stream ends
  OrbitPlan ob;
  u32 wb_groups;   // block writer workgroups (0 = default)
  bool generic_cuts;  // block cuts by the orbit kernels instead of the tiled ones
};

}  // namespace

struct JobPool;

struct luda_ctx {
  int device = 0;
  JobPool* pool = nullptr;  // worker slots of luda_submit, created on first use
  u32 slots = 6;
  hipStream_t s = nullptr;
  u64 budget = 0;
  std::mutex mu;
  DevBuf in_buf;     // uploaded host inputs
  DevBuf a_buf;      // phase A working memory
  DevBuf b_buf;      // phase B working memory
  ScanState scans;   // look-back words of the single-pass scans
  DevStatus* st = nullptr;
  PinnedBuf pin;     // staging for small host<->device copies
  PinnedBuf pin2;
  u32 kw_hint = 0;   // key words that fit the keys seen so far
  // The last two jobs read through phase A: entries and data blocks per
  // input byte, and the decoder's LDS maxima, from which the decoder is
  // launched behind phase A before its counts are back (compact_files).
  struct {
    bool valid;
    double ent_per_byte[2], blk_per_byte[2];
    u32 max_blk, max_ent, max_nr[2];
  } shape{};
  bool decode_ahead = true;  // LUDA_DECODE_AHEAD=0 turns that off
  KTimer T;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool pin2_busy = false;  // a staging kernel that reads pin2 may not have run yet
  // stage_out's completion flag: a word of host memory the kernel stores the
  // sequence number of the staging into, and its workgroup counter
  u64* done_h = nullptr;
  u64* done_d = nullptr;
  u32* done_ctr = nullptr;
  u64 stage_seq = 0;
  std::shared_ptr<DevHeap> heap;  // the family's device memory (working buffers, outputs)

  // output tables of a result; give_out from any thread
  void* take_out(u64 bytes) { return heap->take(bytes); }
  void give_out(void* p) { heap->give(p); }
  // this context (a worker slot) draws on another family's heap
  void use_heap(const std::shared_ptr<DevHeap>& h) {
    heap = h;
    for (DevBuf* b : {&in_buf, &a_buf, &b_buf}) b->heap = h;
  }
  IoState* io = nullptr;  // file I/O stream and staging (io.hip), on first use
  luda_ctx* family = nullptr;  // the context luda_init returned: itself, or a worker slot's owner
  std::vector<luda_tuning_rule> rules;  // luda_set_tuning
};

int ctx_device_of(luda_ctx* c) { return c->device; }
IoState** ctx_io_slot(luda_ctx* c) { return &c->io; }

namespace {

struct Fail {
  int code;
  std::string msg;
};

#define HIPCHK(x)                                                                \
  do {                                                                           \
    hipError_t e__ = (x);                                                        \
    if (e__ != hipSuccess) throw Fail{LUDA_IO_ERROR, std::string("HIP: ") + #x + ": " + hipGetErrorString(e__)}; \
  } while (0)

}  // namespace

void launch_refused(const char* name, hipError_t e) {
  throw Fail{LUDA_IO_ERROR, std::string("HIP refused the launch of ") + name + ": " + hipGetErrorString(e)};
}

void throw_scan_capacity(u64 need, u64 have) {
  throw Fail{LUDA_IO_ERROR, "prefix sum of " + std::to_string(need) + " tiles, words reserved for " +
                                std::to_string(have)};
}

namespace {

int status_code_of(const DevStatus& st) {
  if (st.err & ERR_CORRUPT) return LUDA_CORRUPTION;
  if (st.err & ERR_UNSUPPORTED) return LUDA_NOT_SUPPORTED;
  return LUDA_OK;
}

// Waits for the stream.  The waits of a job are short (its phase A, then its
// phase B), and a blocking wait adds the thread's wake-up to each, so the
// stream is polled for up to 2 ms before the wait blocks.
void stream_sync(luda_ctx* c) {
  const double t0 = now_ms();
  hipError_t e;
  while ((e = hipStreamQuery(c->s)) == hipErrorNotReady && now_ms() - t0 < 2.0) {
  }
  if (e == hipErrorNotReady) e = hipStreamSynchronize(c->s);
  HIPCHK(e);
  c->pin2_busy = false;
}

// Waits for a stage_out's completion flag only: kernels queued behind the
// read-back keep the stream busy while the host reads it.  After 200 ms it
// waits for the stream instead.
void stage_wait_flag(luda_ctx* c) {
  const double t0 = now_ms();
  while (__atomic_load_n(c->done_h, __ATOMIC_ACQUIRE) < c->stage_seq) {
    const double t = now_ms() - t0;
    if (t > 200.0) {
      stream_sync(c);
      return;
    }
    if (t < 1.0) __builtin_ia32_pause();
    else std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
  c->pin2_busy = false;  // the kernels that read pin2 ran before the read-back
}

// Waits for the stream after a stage_out, which ends it.  A worker slot
// first watches the stage_out's completion flag in host memory (no runtime
// call, so the slots' threads do not contend for the runtime while they
// wait: polling the runtime from six threads at once left the device idle
// for about 0.2 ms after each phase of a job); a busy wait for the first
// millisecond, then 20 us naps.  Then a stream_sync, which finds the stream
// done or nearly so.  The flag wait gives up after 200 ms and leaves the
// rest to stream_sync.
void stage_sync(luda_ctx* c) {
  const double t0 = now_ms();
  while (c->family != c && __atomic_load_n(c->done_h, __ATOMIC_ACQUIRE) < c->stage_seq) {
    const double t = now_ms() - t0;
    if (t > 200.0) break;
    if (t < 1.0) __builtin_ia32_pause();
    else std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
  stream_sync(c);
}

// Host bytes to device memory, and with init the status word too, in one
// kernel that reads the pinned staging buffer pin2.  Every job synchronises
// its stream between two staging calls, so pin2 is free again by then; the
// flag only guards against a caller that does not.
void stage_in(luda_ctx* c, void* dst, const void* src, u64 bytes, const DevStatus* init,
              StageFill fill = StageFill{nullptr, 0, 0}) {
  if (c->pin2_busy) stream_sync(c);
  const u64 o_st = align_up(bytes, 64);
  HIPCHK(c->pin2.reserve(o_st + sizeof(DevStatus)));
  if (bytes) memcpy(c->pin2.p, src, bytes);
  if (init) memcpy(c->pin2.p + o_st, init, sizeof(DevStatus));
  launch_stage_in(c->T, c->pin2.d, static_cast<u8*>(dst), bytes,
                  init ? reinterpret_cast<const DevStatus*>(c->pin2.d + o_st) : nullptr, c->st, fill);
  c->pin2_busy = true;
}

// The first launch of a read of tables: the descriptors and the initial
// status go through the pinned staging buffer pin2, which k_files_in reads
// in place (no copy kernel), then footers and the tables' first blocks.
void files_in(luda_ctx* c, FileDesc* d_files, const FileDesc* fdesc, u32 n, const DevStatus& init, FileInfo* d_info,
              u32* d_bbase) {
  if (c->pin2_busy) stream_sync(c);
  const u64 bytes = (u64)n * sizeof(FileDesc);
  const u64 o_st = align_up(bytes, 64);
  HIPCHK(c->pin2.reserve(o_st + sizeof(DevStatus)));
  if (bytes) memcpy(c->pin2.p, fdesc, bytes);
  memcpy(c->pin2.p + o_st, &init, sizeof(DevStatus));
  launch_files_in(c->T, reinterpret_cast<const FileDesc*>(c->pin2.d),
                  reinterpret_cast<const DevStatus*>(c->pin2.d + o_st), d_files, n, d_info, d_bbase, c->st);
  c->pin2_busy = true;
}

// Device bytes (up to three segments) into the pinned buffer pin, at the
// segments' offsets, in one kernel; the caller waits with stage_sync before
// reading pin.  pin must already hold every segment.
void stage_out(luda_ctx* c, StageSeg a, StageSeg b = {}, StageSeg d = {}, bool with_flag = false) {
  for (const StageSeg& g : {a, b, d})
    if (g.bytes && g.off + g.bytes > c->pin.cap) throw Fail{LUDA_IO_ERROR, "read-back staging too small"};
  // only worker slots watch the flag (stage_sync): a context alone on the
  // device finds the stream done sooner by polling the runtime, and skips
  // the flag's system-scope fences
  const bool flag = c->family != c || with_flag;
  launch_stage_out(c->T, c->pin.d, a, b, d,
                   StageDone{flag ? c->done_d : nullptr, c->done_ctr, flag ? ++c->stage_seq : 0});
}

struct EncodeArrays {
  Kept kept;
  EncBufs eb;
  u16* ftab_b;
  u32* ent_b;
  u64* vaddr;  // with LUDA_JOB_KEEP_DECODED: each kept value's address in the output
  CutBufs cut;  // the tiled block cuts' functions and tile entry points
};

// Carves the encoder's arrays for a setup out of an arena (dry or real).
EncodeArrays carve_encode(Arena& a, const EncodeSetup& s, u32 W, bool with_kept, bool with_vaddr = false) {
  EncodeArrays r;
  memset(&r, 0, sizeof(r));
  const u64 m = s.m_cap, nb = s.nbo_cap, nf = s.nf_cap, nw = s.nw_cap;
  if (with_kept) {
    r.kept.K.P = a.take<u64>((u64)W * m);
    r.kept.K.klen = a.take<u32>(m);
    r.kept.K.tag = a.take<u64>(m);
    r.kept.vptr = a.take<u64>(m);
    r.kept.vlen = a.take<u32>(m);
    r.kept.forced = a.take<u32>(m + 1);
    r.kept.K.stride = m;
    r.kept.K.W = W;
  }
  EncBufs& b = r.eb;
  b.shared = a.take<u32>(m);
  b.szr = a.take<u32>(m);
  b.sznr = a.take<u32>(m);
  b.bnext = a.take<u32>(m);
  b.bbytes = a.take<u32>(m);
  b.bs = a.take<u32>(m);
  b.ob_first = a.take<u32>(nb + 1);
  b.ob_n = a.take<u32>(nb + 1);
  b.ob_bytes = a.take<u32>(nb + 1);
  b.ob_forced = a.take<u32>(nb + 1);
  b.tnext = a.take<u16>(nb + 1);
  b.torel = a.take<u64>(nb + 1);
  b.tdend = a.take<u64>(nb + 1);
  b.tnw = a.take<u32>(nb + 1);
  b.tbytes = a.take<u64>(nb / kTableTile + 2);
  b.tforced = a.take<u32>(nb / kTableTile + 2);
  b.ob_file = a.take<u32>(nb + 1);
  b.ob_off = a.take<u64>(nb + 1);
  b.ob_isz = a.take<u64>(nb + 1);
  b.f_first = a.take<u32>(nf + 1);
  b.f_win = a.take<u32>(nf + 1);
  b.f_dend = a.take<u64>(nf + 1);
  b.f_fbytes = a.take<u64>(nf + 1);
  b.f_ibytes = a.take<u64>(nf + 1);
  b.f_fbase = a.take<u64>(nf + 1);
  b.f_off = a.take<u64>(nf + 1);
  b.meta = a.take<FileMeta>(nf + 1);
  b.keys = a.take<u8>((nf + 1) * 2ull * s.ep.kw);
  b.w_keys = a.take<u32>(nw + 1);
  b.w_first = a.take<u32>(nw + 1);
  b.w_file = a.take<u32>(nw + 1);
  b.w_fsz = a.take<u64>(nw + 1);
  r.ftab_b = a.take<u16>(s.ob.table_entries);
  r.ent_b = a.take<u32>(s.ob.entry_entries);
  if (with_vaddr) r.vaddr = a.take<u64>(m);
  const u64 tiles = cut_tiles_for(m);
  r.cut.ftab = a.take<u32>(tiles * kCutCand);
  r.cut.cand = a.take<u32>(tiles);
  r.cut.entry = a.take<u32>(tiles);
  r.cut.base = a.take<u32>(tiles);
  return r;
}

// Whether a job's block cuts start on the orbit kernels: the tiled cuts walk
// each tile's blocks one after another, which for blocks of a few entries
// (large values) is slower than the orbit's pointer jumping.
// LUDA_BLOCK_CUTS=orbit forces the orbit kernels (both give the same cuts).
bool start_generic_cuts(const luda_options& o, u64 entries, u64 bytes) {
  const char* e = getenv("LUDA_BLOCK_CUTS");
  if (e && strcmp(e, "orbit") == 0) return true;
  return entries && (double)o.block_size * (double)entries < 16.0 * (double)std::max<u64>(bytes, 1);
}

// The encoder over the kept stream, counts in st.
void run_encode(luda_ctx* c, const EncodeSetup& s, const EncodeArrays& A, u8* out) {
  KTimer& T = c->T;
  DevStatus* st = c->st;
  const Kept& k = A.kept;
  const EncBufs& b = A.eb;
  const u32* m_dev = &st->n_kept;
  const u32* nbo_dev = &st->n_blocks_out;
  const u32* nf_dev = &st->n_files_out;
  const u32* nw_dev = &st->n_windows;
  ScanState& SC = c->scans;
  HIPCHK(scan_reserve(SC, std::max<u64>(std::max<u64>(s.m_cap, s.nbo_cap), std::max<u64>(s.nf_cap, s.nw_cap)) + 1,
                      c->s));
  HIPCHK(scan_reserve_tiles(SC, (u64)s.nf_cap + 1, c->s));  // the table layout's look-back: one tile a table
  launch_entry_sizes(T, k, m_dev, s.m_cap, b, st);
  if (s.generic_cuts) {
    launch_bnext(T, k, m_dev, s.m_cap, s.ep, b, st);
    run_orbit(T, s.ob, b.bnext, m_dev, &st->blk_step_max, b.bs, A.ftab_b, A.ent_b, st);
    scan_blocks(T, SC, k, m_dev, s.m_cap, b, s.nbo_cap, st);
  } else {
    run_block_cuts(T, k, m_dev, s.m_cap, s.ep, b, A.cut, s.nbo_cap, st);
  }
  launch_cut_tables(T, nbo_dev, s.nbo_cap, s.ep, b, s.nf_cap, s.nw_cap, s.size_cuts, st);
  launch_table_layout(T, SC, k, nf_dev, s.nf_cap, s.ep, b, st);
  launch_write_blocks(T, k, nbo_dev, s.nbo_cap, s.ep, b, out, s.out_cap, st, s.wb_groups);
  if (s.ep.bloom_bits) launch_write_filters(T, k, nw_dev, s.nw_cap, nf_dev, s.ep, b, out, s.out_cap, st);
  launch_write_tail(T, k, nf_dev, s.nf_cap, s.ep, b, out, s.out_cap, st);
  if (A.vaddr) launch_value_addrs(T, k, nbo_dev, s.nbo_cap, s.ep, b, out, A.vaddr);
}

// Orbit chunk for successor steps of at most `step` nodes: a chunk must hold
// one whole step (an orbit cannot jump over a chunk); twice that when it fits
// keeps the per-chunk functions small relative to the chunk.
u32 orbit_chunk(u64 step, u32 min_chunk) {
  const u64 two = next_pow2(2 * step);
  if (two <= kMaxChunk) return std::max<u32>(min_chunk, (u32)two);
  const u64 one = next_pow2(step);
  return one > kMaxChunk ? (u32)std::min<u64>(one, 1u << 31) : std::max<u32>(min_chunk, (u32)one);
}

// Capacities and orbit plans for an encoder run over at most m entries.
EncodeSetup plan_encode(const luda_options& o, u32 W, u64 m, u64 out_cap, u64 extra_files, u64 nf_cap_min,
                        u64 nw_cap_min, u64 nbo_cap_min) {
  EncodeSetup s;
  memset(&s.ep, 0, sizeof(s.ep));
  s.wb_groups = 0;
  s.ep.block_size = o.block_size;
  s.ep.max_file_size = o.max_file_size;
  s.ep.interval = o.block_restart_interval;
  s.ep.bloom_bits = o.bloom_bits_per_key;
  u64 k = (u64)(o.bloom_bits_per_key * 0.69);  // util/bloom.cc:21
  if (k < 1) k = 1;
  if (k > 30) k = 30;
  s.ep.bloom_k = (u32)k;
  s.ep.kw = 8 * W + 8;
  s.m_cap = (u32)std::max<u64>(m, 1);
  s.out_cap = out_cap;
  s.size_cuts = o.max_file_size < out_cap;
  u64 nf = s.size_cuts ? out_cap / std::max<u64>(o.max_file_size, 1) + extra_files + 2 : extra_files + 2;
  nf = std::max<u64>(nf, nf_cap_min);
  // A block closed for size holds at least block_size bytes; the others end
  // a table (<= nf), precede a ShouldStopBefore cut (<= extra_files) or end
  // the stream.  A count over this bound is caught on the device and retried.
  const u64 nbo = out_cap / o.block_size + nf + extra_files + 2;
  s.nbo_cap = (u32)std::min<u64>(s.m_cap, std::max<u64>(nbo, nbo_cap_min));
  s.nf_cap = (u32)std::min<u64>(nf, s.nbo_cap);
  u64 nw = o.bloom_bits_per_key ? out_cap / 2048 + s.nf_cap + 2 : 1;
  s.nw_cap = (u32)std::max<u64>(nw, nw_cap_min);
  // blocks: an entry is at least 4 bytes and 16 of them add a 4-byte restart
  const u64 sb = o.block_size / 4 + 2;
  const u32 Cb = orbit_chunk(sb, 2048);
  const u32 Wb = (u32)std::min<u64>(sb, Cb);
  s.ob = plan_orbit(s.m_cap, Cb, Wb, std::max<u32>(2, std::min<u32>(32, 16384 / Wb)));
  if (Cb > kMaxChunk) throw Fail{LUDA_NOT_SUPPORTED, "block_size too large for the device block cut"};
  return s;
}

void validate_options(const luda_options& o) {
  if (o.compression != 0) throw Fail{LUDA_NOT_SUPPORTED, "only kNoCompression is supported"};
  if (o.block_restart_interval < 1) throw Fail{LUDA_INVALID_ARGUMENT, "block_restart_interval < 1"};
  if (o.block_size < 1 || o.block_size > kMaxBlockSize)
    throw Fail{LUDA_NOT_SUPPORTED, "block_size outside what the device encoder handles"};
  if (o.max_file_size < 1) throw Fail{LUDA_INVALID_ARGUMENT, "max_file_size < 1"};
  if (o.bloom_bits_per_key > 1000) throw Fail{LUDA_NOT_SUPPORTED, "bloom_bits_per_key too large"};
}

// LUDA_KERNEL_TIMING=1 times every kernel with HIP events on the engine's
// stream; LUDA_KERNEL_TIMING=name[,name...] times only those kernels.
// Also LUDA_DECODE_AHEAD=0: no decoder launched behind phase A's read-back.
void init_ctx_timer(luda_ctx* c) {
  const char* da = getenv("LUDA_DECODE_AHEAD");
  c->decode_ahead = !(da && strcmp(da, "0") == 0);
  const char* e = getenv("LUDA_KERNEL_TIMING");
  c->T.on = e && *e && *e != '0';
  c->T.only.clear();
  if (c->T.on && strcmp(e, "1") != 0) {
    std::string s(e);
    size_t p = 0;
    while (p <= s.size()) {
      size_t q = s.find(',', p);
      if (q == std::string::npos) q = s.size();
      if (q > p) c->T.only.push_back(s.substr(p, q - p));
      p = q + 1;
    }
  }
  c->T.reset();
}

// Algorithmic HBM bytes of one launch of the kernels worth a roofline
// (each input byte read once, each output byte written once):
//   decode        data blocks read + key planes, tags, value refs written
//   merge         per round: indices read and written + the keys compared once
//   write_blocks  key and value bytes read + data blocks written
//   write_filters user keys read + filter bytes written (approx. 1.25 B/key)
// Algorithmic HBM bytes of all launches of a kernel in one call: what it has
// to read and write at least, from the job's counts (n entries in, m kept, W
// key words, soa = bytes of one entry's planes and value reference).
u64 kernel_bytes(const std::string& name, const luda_stats& s, u32 W, u64 rounds) {
  const u64 n = s.entries_in, m = s.entries_kept;
  const u64 soa = 8ull * W + 4 + 8 + 8 + 4;
  const u64 tail = s.bytes_written > s.data_block_bytes ? s.bytes_written - s.data_block_bytes : 0;
  if (name == "decode") return s.bytes_read + n * soa;               // input tables in, planes out
  if (name == "merge") return (rounds ? rounds - 1 : 0) * n * (4 + 4 + 8ull * W + 12);  // per round: perm in/out + keys
  if (name == "merge_select") return n * (8ull * W + 12 + 12) + m * soa;  // keys and value refs in, kept stream out
  if (name == "drop") return n * (4 + 4 + 8 + 8ull * W + 4);
  if (name == "gather") return n * 8 + 2 * m * soa;
  if (name == "entry_sizes") return m * (2 * (8ull * W + 12) + 12);  // two keys in, three sizes out
  if (name == "bnext") return m * (4 + 4 + 4 + 4 + 4);               // sizes + forced in, successor + bytes out
  if (name == "write_blocks") return 2 * s.data_block_bytes;          // entries in, blocks out
  if (name == "write_filters") return m * (8ull * W + 4) + m * 2;     // keys in, ~10 bits per key out
  if (name == "write_tail") return 2 * tail;                          // filter data in (CRC), tail blocks out
  if (name == "index") return s.blocks_in * 56;  // index entries and restart arrays in, handles and counts out
  return 0;
}

void fill_kernel_stats(luda_ctx* c, luda_stats* s, u32 W, u64 rounds) {
  if (!c->T.on) return;
  std::map<std::string, std::pair<double, u64>> acc;
  for (auto& r : c->T.recs) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      auto& x = acc[r.name];
      x.first += ms;
      x.second++;
    }
  }
  double best = -1;
  for (auto& kv : acc) {
    if (kv.second.first > best) {
      best = kv.second.first;
      s->top_kernel_ms = kv.second.first;
      s->top_kernel_launches = kv.second.second;
      snprintf(s->top_kernel, sizeof(s->top_kernel), "%s", kv.first.c_str());
    }
  }
  if (best >= 0) s->top_kernel_bytes = kernel_bytes(s->top_kernel, *s, W, rounds);
}

// Reads the outputs back and builds the result.
// hmeta / hkeys: host copies of the table metadata and boundary keys, read
// back with the job's status (nf <= the capacity they were read at).
luda_result* finish_result(luda_ctx* c, const EncodeSetup& s, const EncodeArrays& A, void* out, const DevStatus& st,
                           u32 flags, luda_stats stats, const FileMeta* hmeta, const u8* hkeys) {
  (void)A;
  const u32 nf = st.n_files_out;
  luda_result* r = (luda_result*)calloc(1, sizeof(luda_result));
  r->num_outputs = nf;
  r->outputs = (luda_output*)calloc(nf ? nf : 1, sizeof(luda_output));
  std::vector<FileMeta> meta(hmeta, hmeta + nf);
  std::vector<u8> keys(hkeys, hkeys + (u64)nf * 2 * s.ep.kw);
  u64 kbytes = 0;
  for (u32 i = 0; i < nf; i++) kbytes += meta[i].slen + meta[i].llen;
  r->keys = (uint8_t*)malloc(kbytes ? kbytes : 1);
  u64 ko = 0;
  for (u32 i = 0; i < nf; i++) {
    luda_output& o = r->outputs[i];
    o.offset = meta[i].offset;
    o.file_size = meta[i].size;
    o.num_entries = meta[i].nent;
    o.smallest_offset = ko;
    o.smallest_size = meta[i].slen;
    memcpy(r->keys + ko, &keys[(u64)i * 2 * s.ep.kw], meta[i].slen);
    ko += meta[i].slen;
    o.largest_offset = ko;
    o.largest_size = meta[i].llen;
    memcpy(r->keys + ko, &keys[(u64)i * 2 * s.ep.kw + s.ep.kw], meta[i].llen);
    ko += meta[i].llen;
  }
  r->total_bytes = st.out_bytes;
  r->device_data = out;
  if (!(flags & LUDA_JOB_KEEP_ON_DEVICE)) {
    const double t = now_ms();
    void* h = nullptr;
    if (hipHostMalloc(&h, std::max<u64>(st.out_bytes, 1), hipHostMallocDefault) != hipSuccess) {
      free(r->outputs);
      free(r->keys);
      free(r);
      throw Fail{LUDA_IO_ERROR, "pinned host allocation for the output failed"};
    }
    if (st.out_bytes) HIPCHK(hipMemcpyAsync(h, out, st.out_bytes, hipMemcpyDeviceToHost, c->s));
    stream_sync(c);
    r->data = (uint8_t*)h;
    stats.d2h_ms = now_ms() - t;
  }
  stats.bytes_written = st.out_bytes;
  stats.entries_kept = st.n_kept;
  stats.blocks_out = st.n_blocks_out;
  stats.data_block_bytes = st.data_bytes;
  r->stats = stats;
  return r;
}

// ---------------------------------------------------------------------------
// Resident decoded entries (LUDA_JOB_KEEP_DECODED).  A result's kept stream
// (planes, tags, value lengths) and its values' addresses in the output
// tables are copied to a buffer the result owns; each output table is
// registered by its device address, so that a later job whose inputs are all
// registered tables gathers their entries instead of decoding them.

struct DecodedSet {
  const luda_ctx* family;
  std::shared_ptr<DevHeap> heap;  // buf's
  void* buf;
  u32 W;
  u64 m;  // entries (the stride of every array)
  u64* P;
  u32* klen;
  u64* tag;
  u64* vptr;
  u32* vlen;
};
struct ResidentTable {
  DecodedSet* d;
  u64 first, count, size;
};
// Tables are registered for the context family that wrote them (a context
// and its worker slots): another context's jobs decode them as usual.
struct ResidentKey {
  const luda_ctx* family;
  u64 addr;
  bool operator<(const ResidentKey& o) const { return family != o.family ? family < o.family : addr < o.addr; }
};
std::mutex g_resident_mu;
std::map<ResidentKey, ResidentTable> g_resident;
std::map<const luda_result*, DecodedSet*> g_decoded_of;

// A result this context just built, released on a failure after it was
// made (its device buffer goes back to the context's pool).
void free_result_now(luda_ctx* c, luda_result* r) {
  if (r->device_data) c->give_out(r->device_data);
  if (r->data) hipHostFree(r->data);
  free(r->outputs);
  free(r->keys);
  free(r);
}

// After the job's last synchronisation: its m kept entries copied out of the
// context's working memory, and its tables registered.
void keep_decoded(luda_ctx* c, const EncodeArrays& A, u32 W, u64 m, luda_result* r) {
  DecodedSet* d = new DecodedSet;
  d->family = c->family;
  d->heap = c->heap;
  d->W = W;
  d->m = m;
  const u64 mm = std::max<u64>(m, 1);
  const u64 bytes = mm * (8ull * W + 4 + 8 + 8 + 4) + 5 * kAlign;
  d->buf = d->heap->take(bytes);
  if (!d->buf) {
    delete d;
    throw Fail{LUDA_IO_ERROR, "device memory for the resident decoded entries"};
  }
  Arena a;
  a.base = static_cast<u8*>(d->buf);
  a.cap = bytes;
  d->P = a.take<u64>((u64)W * mm);
  d->klen = a.take<u32>(mm);
  d->tag = a.take<u64>(mm);
  d->vptr = a.take<u64>(mm);
  d->vlen = a.take<u32>(mm);
  const Kept& k = A.kept;
  hipError_t e = hipSuccess;
  for (u32 w = 0; w < W && m && e == hipSuccess; w++)
    e = hipMemcpyAsync(d->P + (u64)w * m, k.K.P + (u64)w * k.K.stride, m * 8, hipMemcpyDeviceToDevice, c->s);
  if (m && e == hipSuccess) e = hipMemcpyAsync(d->klen, k.K.klen, m * 4, hipMemcpyDeviceToDevice, c->s);
  if (m && e == hipSuccess) e = hipMemcpyAsync(d->tag, k.K.tag, m * 8, hipMemcpyDeviceToDevice, c->s);
  if (m && e == hipSuccess) e = hipMemcpyAsync(d->vptr, A.vaddr, m * 8, hipMemcpyDeviceToDevice, c->s);
  if (m && e == hipSuccess) e = hipMemcpyAsync(d->vlen, k.vlen, m * 4, hipMemcpyDeviceToDevice, c->s);
  if (e == hipSuccess) e = hipStreamSynchronize(c->s);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(c->s);  // no copy still writes the block
    d->heap->give(d->buf);
    delete d;
    throw Fail{LUDA_IO_ERROR, std::string("resident decoded entries: ") + hipGetErrorString(e)};
  }
  std::lock_guard<std::mutex> lk(g_resident_mu);
  u64 first = 0;
  for (u32 i = 0; i < r->num_outputs; i++) {
    const luda_output& o = r->outputs[i];
    g_resident[ResidentKey{c->family, (u64)r->device_data + o.offset}] = ResidentTable{d, first, o.num_entries, o.file_size};
    first += o.num_entries;
  }
  g_decoded_of[r] = d;
}

void forget_decoded(const luda_result* r) {
  DecodedSet* d = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_resident_mu);
    auto it = g_decoded_of.find(r);
    if (it == g_decoded_of.end()) return;
    d = it->second;
    g_decoded_of.erase(it);
    for (u32 i = 0; i < r->num_outputs; i++) g_resident.erase(ResidentKey{d->family, (u64)r->device_data + r->outputs[i].offset});
  }
  d->heap->give(d->buf);
  delete d;
}

// Every input table registered for this context family (same size): their
// sources, in input order.
bool find_resident(const luda_ctx* family, const std::vector<const luda_file*>& files, std::vector<DecodedSrc>* out) {
  std::lock_guard<std::mutex> lk(g_resident_mu);
  out->clear();
  for (const luda_file* f : files) {
    auto it = g_resident.find(ResidentKey{family, (u64)f->data});
    if (it == g_resident.end() || it->second.size != f->size) return false;
    const ResidentTable& t = it->second;
    DecodedSrc s;
    memset(&s, 0, sizeof(s));
    s.P = t.d->P;
    s.stride = t.d->m;
    s.W = t.d->W;
    s.klen = t.d->klen;
    s.tag = t.d->tag;
    s.vptr = t.d->vptr;
    s.vlen = t.d->vlen;
    s.first = t.first;
    s.count = t.count;
    out->push_back(s);
  }
  return !files.empty();
}

// The context's tuning for a job of `bytes` input bytes and user keys of at
// most key_bytes (a multiple of 8): the last matching rule.
luda_tuning pick_tuning(const luda_ctx* c, u64 bytes, u32 key_bytes) {
  luda_tuning t;
  memset(&t, 0, sizeof(t));
  for (const luda_tuning_rule& r : c->rules)
    if (bytes >= r.min_bytes && bytes < r.max_bytes && key_bytes >= r.min_key_bytes && key_bytes <= r.max_key_bytes)
      t = r.tuning;
  return t;
}

// ---------------------------------------------------------------------------
// luda_compact

struct Child {
  u32 f0, f1;  // files [f0, f1)
};

luda_result* compact(luda_ctx* c, const luda_job* job) {
  const double t_call = now_ms();
  luda_stats stats;
  memset(&stats, 0, sizeof(stats));
  validate_options(job->options);
  if (job->level < 0 || job->level >= 6) throw Fail{LUDA_INVALID_ARGUMENT, "level out of range"};
  // Inputs in MakeInputIterator order (db/version_set.cc:1237-1268).
  std::vector<const luda_file*> files;
  std::vector<Child> children;
  for (int which = 0; which < 2; which++) {
    const u32 n = job->num_inputs[which];
    if (!n) continue;
    const u32 f0 = (u32)files.size();
    for (u32 i = 0; i < n; i++) files.push_back(&job->inputs[which][i]);
    if (job->level + which == 0) {
      for (u32 i = 0; i < n; i++) children.push_back({f0 + i, f0 + i + 1});
    } else {
      children.push_back({f0, f0 + n});
    }
  }
  const u32 nfiles = (u32)files.size();
  u64 bytes_in = 0;
  for (auto* f : files) bytes_in += f->size;
  stats.bytes_read = bytes_in;
  // Grandparents: sorted internal keys (Compaction::grandparents_ come from a
  // sorted level); deeper levels: Version::files_ order, sorted by key.
  for (u32 i = 0; i < job->num_grandparents; i++) {
    const luda_slice& k = job->grandparents[i].largest;
    if (k.size < 8) throw Fail{LUDA_NOT_SUPPORTED, "grandparent key shorter than 8 bytes"};
    if (i > 0) {
      const luda_slice& p = job->grandparents[i - 1].largest;
      if (ikey_cmp_host(p.data, p.size, k.data, k.size) > 0) throw Fail{LUDA_NOT_SUPPORTED, "grandparents not sorted"};
    }
  }
  std::vector<std::vector<u32>> lvl_files(8);
  for (u32 i = 0; i < job->num_deeper; i++) {
    const int L = job->deeper[i].level;
    if (L >= job->level + 2 && L < 7) lvl_files[L].push_back(i);
  }
  for (int L = 0; L < 7; L++)
    for (size_t j = 1; j < lvl_files[L].size(); j++) {
      const luda_level_file& a = job->deeper[lvl_files[L][j - 1]];
      const luda_level_file& b = job->deeper[lvl_files[L][j]];
      if (bytewise_cmp(a.largest_user_key.data, a.largest_user_key.size, b.largest_user_key.data,
                       b.largest_user_key.size) > 0)
        throw Fail{LUDA_NOT_SUPPORTED, "deeper level files not sorted"};
    }
  if (nfiles == 0) {
    luda_result* r = (luda_result*)calloc(1, sizeof(luda_result));
    r->outputs = (luda_output*)calloc(1, sizeof(luda_output));
    r->keys = (uint8_t*)malloc(1);
    r->stats = stats;
    r->stats.attempts = 1;
    r->stats.end_clock_ms = now_ms();
    return r;
  }

  HIPCHK(c->pin.reserve(1024 + 4ull * (nfiles + 1) + sizeof(DevStatus)));
  init_ctx_timer(c);
  // ---- inputs resident in HBM
  std::vector<FileDesc> fdesc(nfiles);
  if (job->flags & LUDA_JOB_INPUTS_ON_DEVICE) {
    for (u32 i = 0; i < nfiles; i++) fdesc[i] = {files[i]->data, files[i]->size};
  } else {
    const double t = now_ms();
    u64 need = 0;
    for (auto* f : files) need += align_up(f->size + 16, 256);
    HIPCHK(c->in_buf.reserve(need));
    u64 off = 0;
    for (u32 i = 0; i < nfiles; i++) {
      HIPCHK(hipMemcpyAsync(c->in_buf.p + off, files[i]->data, files[i]->size, hipMemcpyHostToDevice, c->s));
      fdesc[i] = {c->in_buf.p + off, files[i]->size};
      off += align_up(files[i]->size + 16, 256);
    }
    stream_sync(c);
    stats.h2d_ms = now_ms() - t;
  }
  HIPCHK(hipEventRecord(c->ev0, c->s));

  // Inputs whose decoded entries an earlier call left resident
  // (LUDA_JOB_KEEP_DECODED): when every input is one, phase A and the decode
  // are skipped and their entries gathered instead.
  std::vector<DecodedSrc> dsrc;
  const bool resident = (job->flags & LUDA_JOB_INPUTS_ON_DEVICE) && find_resident(c->family, files, &dsrc);
  FileDesc* d_files = nullptr;
  FileInfo* d_info = nullptr;
  u32 *d_bbase = nullptr, *d_nent = nullptr, *d_ebase = nullptr, *d_febase = nullptr;
  BlockDesc* d_blks = nullptr;
  DevStatus stA;
  memset(&stA, 0, sizeof(stA));
  std::vector<u32> febase(nfiles + 1);
  KTimer& T = c->T;
  // the decoder launched behind phase A (below): its arrays, in phase A's
  // memory, and whether phase A's counts turned out to fit it
  Planes gK{};
  u64* g_vptr = nullptr;
  u32* g_vlen = nullptr;
  u64 n_guess = 0;
  const u32 W_guess = std::max<u32>(1, c->kw_hint);
  bool guessed = false;
  if (resident) {
    u64 e = 0;
    u32 wmax = 1;
    for (u32 i = 0; i < nfiles; i++) {
      febase[i] = (u32)e;
      dsrc[i].dst = e;
      e += dsrc[i].count;
      wmax = std::max<u32>(wmax, dsrc[i].W);
    }
    if (e > 0xffffffffull / 2) throw Fail{LUDA_NOT_SUPPORTED, "inputs too large"};
    febase[nfiles] = (u32)e;
    stA.n_entries = (u32)e;
    stA.max_rulen = 8 * wmax;
  } else {
    // ---- phase A: footers, index blocks, restart arrays
    u64 nb_cap = 0;
    for (auto* f : files) nb_cap += f->size / 41 + 1;
    if (nb_cap > 0xffffffffull / 2) throw Fail{LUDA_NOT_SUPPORTED, "inputs too large"};
    // The decoder goes behind phase A's read-back, shaped like the context's
    // last two jobs (entries per input byte with a quarter more, their
    // largest restart arrays, the key width seen so far), before the host
    // has read phase A's counts: it decodes only when they fit that launch
    // (k_decode_entries checks them first), and the host, which finds the
    // same from the counts, then plans phase B while it runs.  Otherwise it
    // is launched again once the counts are known, as without a guess.
    const auto& sh = c->shape;
    const u32 g_max_nr = std::max(sh.max_nr[0], sh.max_nr[1]);
    bool guess = c->decode_ahead && sh.valid && W_guess <= 8 && decode_lds_bytes(sh.max_blk, g_max_nr) <= 65536;
    if (guess) {
      n_guess = (u64)((double)bytes_in * std::max(sh.ent_per_byte[0], sh.ent_per_byte[1]) * 1.25) + 1024;
      guess = n_guess <= 0x7fffffffull && n_guess * (8ull * W_guess + 32) <= c->budget / 4;
    }
    auto carve_a = [&](Arena& a, FileDesc*& d_files, FileInfo*& d_info, u32*& d_bbase, BlockDesc*& d_blks,
                       u32*& d_nent, u32*& d_ebase, u32*& d_febase) {
      d_files = a.take<FileDesc>(nfiles);
      d_info = a.take<FileInfo>(nfiles);
      d_bbase = a.take<u32>(nfiles + 1);
      d_blks = a.take<BlockDesc>(nb_cap);
      d_nent = a.take<u32>(nb_cap + 1);
      d_ebase = a.take<u32>(nb_cap + 1);
      d_febase = a.take<u32>(nfiles + 1);
      if (guess) {
        gK.P = a.take<u64>((u64)W_guess * n_guess);
        gK.klen = a.take<u32>(n_guess);
        gK.tag = a.take<u64>(n_guess);
        gK.stride = n_guess;
        gK.W = W_guess;
        g_vptr = a.take<u64>(n_guess);
        g_vlen = a.take<u32>(n_guess);
      }
    };
    {
      Arena dry;
      dry.dry = true;
      carve_a(dry, d_files, d_info, d_bbase, d_blks, d_nent, d_ebase, d_febase);
      HIPCHK(c->a_buf.reserve(dry.used));
      Arena a;
      a.base = c->a_buf.p;
      a.cap = c->a_buf.cap;
      carve_a(a, d_files, d_info, d_bbase, d_blks, d_nent, d_ebase, d_febase);
    }
    DevStatus st0;
    memset(&st0, 0, sizeof(st0));
    HIPCHK(scan_reserve_tiles(c->scans, index_scan_tiles((u32)nb_cap), c->s));
    files_in(c, d_files, fdesc.data(), nfiles, st0, d_info, d_bbase);
    // sets n_blocks_in, each block's and table's first entry, n_entries
    launch_index_entries(T, c->scans, d_files, d_info, d_bbase, nfiles, (u32)nb_cap, d_blks, d_nent, d_ebase, d_febase,
                         c->st);
    HIPCHK(hipGetLastError());
    stage_out(c, StageSeg{c->st, 0, sizeof(DevStatus)}, StageSeg{d_febase, 1024, (nfiles + 1) * 4ull}, StageSeg{},
              guess);
    if (guess) {
      const u64 nb_guess =
          std::min<u64>(nb_cap, (u64)((double)bytes_in * std::max(sh.blk_per_byte[0], sh.blk_per_byte[1]) * 1.25) + 64);
      launch_decode(T, d_files, d_blks, (u32)nb_guess, d_ebase, gK, g_vptr, g_vlen, c->st, sh.max_blk, g_max_nr,
                    sh.max_ent, pick_tuning(c, bytes_in, 8 * W_guess).decode_groups,
                    DecodeGuard{1, (u32)n_guess, 8 * W_guess});
      HIPCHK(hipGetLastError());
      stage_wait_flag(c);
    } else {
      stage_sync(c);
    }
    memcpy(&stA, c->pin.p, sizeof(DevStatus));
    memcpy(febase.data(), c->pin.p + 1024, (nfiles + 1) * 4);
    if (status_code_of(stA) != LUDA_OK)
      throw Fail{status_code_of(stA), std::string("input rejected: ") + site_name(stA.site)};
    if (stA.n_blocks_in > nb_cap) throw Fail{LUDA_CORRUPTION, "more data blocks than the tables can hold"};
    // the test k_decode_entries made
    guessed = guess && stA.err == 0 && stA.site == 0 && stA.max_ulen == 0 && stA.n_entries <= n_guess &&
              stA.max_rulen <= 8 * W_guess && stA.max_nr <= g_max_nr;
    auto& shw = c->shape;
    shw.ent_per_byte[1] = shw.ent_per_byte[0];
    shw.blk_per_byte[1] = shw.blk_per_byte[0];
    shw.max_nr[1] = shw.max_nr[0];
    shw.ent_per_byte[0] = (double)stA.n_entries / (double)std::max<u64>(bytes_in, 1);
    shw.blk_per_byte[0] = (double)stA.n_blocks_in / (double)std::max<u64>(bytes_in, 1);
    shw.max_nr[0] = stA.max_nr;
    shw.max_blk = stA.max_blk;
    shw.max_ent = stA.max_ent;
    shw.valid = true;
  }
  const u32 nb = stA.n_blocks_in;
  const u32 n = stA.n_entries;
  stats.blocks_in = nb;
  stats.entries_in = n;

  // ---- phase B
  u32 W = std::max<u32>(1, (std::max<u32>(stA.max_rulen, 8 * c->kw_hint) + 7) / 8);
  u64 out_cap = bytes_in + bytes_in / 4 + (1 << 20);
  u64 nf_min = 0, nw_min = 0, nbo_min = 0;
  // Runs to merge: children's entry ranges, empty ones left out.
  std::vector<u32> cstart;
  std::vector<std::pair<u32, u32>> runs;
  for (auto& ch : children) {
    const u32 e0 = febase[ch.f0], e1 = febase[ch.f1];
    cstart.push_back(e0);
    if (e1 > e0) runs.push_back({e0, e1});
  }
  cstart.push_back(n);
  // merge rounds, tiled for key planes of W words (planned again when W grows)
  std::vector<std::vector<MergePair>> rounds;
  auto plan_rounds = [&](u32 Wk) {
    rounds.clear();
    auto cur = runs;
    const u32 items = merge_items_per_tile(Wk);
    while (cur.size() > 1) {
      std::vector<MergePair> prs;
      std::vector<std::pair<u32, u32>> nxt;
      u32 tiles = 0;
      for (size_t i = 0; i < cur.size(); i += 2) {
        MergePair p;
        p.a0 = cur[i].first;
        p.a1 = cur[i].second;
        p.b1 = i + 1 < cur.size() ? cur[i + 1].second : cur[i].second;
        p.tile0 = tiles;
        tiles += (p.b1 - p.a0 + items - 1) / items;
        prs.push_back(p);
        nxt.push_back({p.a0, p.b1});
      }
      rounds.push_back(prs);
      cur = nxt;
    }
  };
  const u32 ngp = job->num_grandparents;
  u32 ndeep = 0;
  for (int L = 0; L < 7; L++) ndeep += (u32)lvl_files[L].size();
  bool generic_cuts = start_generic_cuts(job->options, n, bytes_in);

  for (u32 attempt = 1;; attempt++) {
    if (W > kMaxKeyWords) throw Fail{LUDA_NOT_SUPPORTED, "user keys longer than the device key planes"};
    plan_rounds(W);
    EncodeSetup es = plan_encode(job->options, W, n, out_cap, ngp, nf_min, nw_min, nbo_min);
    const luda_tuning tune = pick_tuning(c, bytes_in, 8 * W);
    es.wb_groups = tune.write_blocks_groups;
    es.generic_cuts = generic_cuts;
    // host-side tables (grandparents, deeper levels, children, merge pairs)
    // staged in one blob that a single copy uploads
    std::vector<u64> gw((u64)std::max<u32>(ngp, 1) * W, 0), gt(std::max<u32>(ngp, 1), 0), gpre(ngp + 1, 0);
    std::vector<u32> gl(std::max<u32>(ngp, 1), 0);
    for (u32 i = 0; i < ngp; i++) {
      const luda_slice& k = job->grandparents[i].largest;
      to_words(k.data, k.size - 8, W, &gw[(u64)i * W]);
      gl[i] = (u32)(k.size - 8);
      memcpy(&gt[i], k.data + k.size - 8, 8);
      gpre[i + 1] = gpre[i] + job->grandparents[i].file_size;
    }
    std::vector<u64> dw((u64)std::max<u32>(ndeep, 1) * 2 * W, 0);
    std::vector<u32> dl(std::max<u32>(ndeep, 1) * 2, 0);
    DropParams dp;
    memset(&dp, 0, sizeof(dp));
    {
      u32 di = 0;
      for (int L = 0; L < 7; L++) {
        dp.lvl_begin[L] = di;
        for (u32 idx : lvl_files[L]) {
          const luda_level_file& f = job->deeper[idx];
          to_words(f.smallest_user_key.data, f.smallest_user_key.size, W, &dw[(2ull * di) * W]);
          to_words(f.largest_user_key.data, f.largest_user_key.size, W, &dw[(2ull * di + 1) * W]);
          dl[2 * di] = (u32)f.smallest_user_key.size;
          dl[2 * di + 1] = (u32)f.largest_user_key.size;
          di++;
        }
        dp.lvl_end[L] = di;
      }
    }
    std::vector<MergePair> allp;
    for (auto& r : rounds) allp.insert(allp.end(), r.begin(), r.end());
    // the last round's pair again, at the end, for the fused last round (a
    // single run is merged with an empty one)
    MergePair fin{0, 0, 0, 0};
    if (!rounds.empty()) fin = rounds.back().front();
    else if (!runs.empty()) fin = MergePair{runs[0].first, runs[0].second, runs[0].second, 0};
    allp.push_back(fin);
    std::vector<u8> blob;
    auto put = [&](const void* p, u64 bytes) {
      const u64 off = align_up(blob.size(), 16);
      blob.resize(off + bytes);
      if (bytes) memcpy(blob.data() + off, p, bytes);
      return off;
    };
    const u64 o_gw = put(gw.data(), gw.size() * 8), o_gl = put(gl.data(), gl.size() * 4),
              o_gt = put(gt.data(), gt.size() * 8), o_gp = put(gpre.data(), gpre.size() * 8),
              o_dw = put(dw.data(), dw.size() * 8), o_dl = put(dl.data(), dl.size() * 4),
              o_cs = put(cstart.data(), cstart.size() * 4),
              o_mp = put(allp.data(), allp.size() * sizeof(MergePair)),
              o_ds = put(dsrc.data(), dsrc.size() * sizeof(DecodedSrc));
    // carve phase B
    Planes K;
    u64 *vptr, *gp_words, *gp_tag, *gp_prefix, *deep_words;
    u32 *vlen, *permA, *permB, *keep, *kpos, *gcnt, *events, *gp_len, *deep_len, *d_cstart, *splits;
    MergePair* d_pairs;
    u8* d_tab;
    EncodeArrays EA;
    // the first attempt takes the decoded arrays of the decoder launched
    // behind phase A when the counts fitted it (same key width)
    const bool use_guess = guessed && attempt == 1 && W == W_guess;
    auto carve_b = [&](Arena& a) {
      if (use_guess) {
        K = gK;
        vptr = g_vptr;
        vlen = g_vlen;
      } else {
        K.P = a.take<u64>((u64)W * n);
        K.klen = a.take<u32>(n);
        K.tag = a.take<u64>(n);
        K.stride = n;
        K.W = W;
        vptr = a.take<u64>(n);
        vlen = a.take<u32>(n);
      }
      permA = a.take<u32>(n);
      permB = a.take<u32>(n);
      keep = a.take<u32>(n);
      kpos = a.take<u32>(n);
      gcnt = a.take<u32>(ngp ? n : 1);
      events = a.take<u32>(ngp + 2);
      splits = a.take<u32>((u64)n / merge_items_per_tile(W) + runs.size() + 2);
      d_tab = a.take<u8>(blob.size());
      EA = carve_encode(a, es, W, true, (job->flags & LUDA_JOB_KEEP_DECODED) != 0);
    };
    {
      Arena dry;
      dry.dry = true;
      carve_b(dry);
      if (dry.used > c->budget) throw Fail{LUDA_NOT_SUPPORTED, "job exceeds the HBM budget"};
      HIPCHK(c->b_buf.reserve(dry.used));
      Arena a;
      a.base = c->b_buf.p;
      a.cap = c->b_buf.cap;
      carve_b(a);
    }
    gp_words = reinterpret_cast<u64*>(d_tab + o_gw);
    gp_len = reinterpret_cast<u32*>(d_tab + o_gl);
    gp_tag = reinterpret_cast<u64*>(d_tab + o_gt);
    gp_prefix = reinterpret_cast<u64*>(d_tab + o_gp);
    deep_words = reinterpret_cast<u64*>(d_tab + o_dw);
    deep_len = reinterpret_cast<u32*>(d_tab + o_dl);
    d_cstart = reinterpret_cast<u32*>(d_tab + o_cs);
    d_pairs = reinterpret_cast<MergePair*>(d_tab + o_mp);
    const DecodedSrc* d_dsrc = reinterpret_cast<const DecodedSrc*>(d_tab + o_ds);
    void* out = c->take_out(std::max<u64>(out_cap, 1));
    if (!out) throw Fail{LUDA_IO_ERROR, "output allocation failed"};
    {

      dp.S = job->smallest_snapshot;
      dp.limit = 10 * job->options.max_file_size;  // MaxGrandParentOverlapBytes, db/version_set.cc:31-33
      dp.ngp = ngp;
      dp.ndeep = ndeep;
      dp.gp_words = gp_words;
      dp.gp_len = gp_len;
      dp.gp_tag = gp_tag;
      dp.gp_prefix = gp_prefix;
      dp.deep_words = deep_words;
      dp.deep_len = deep_len;
      dp.first_level = (u32)job->level + 2;

      // status for this attempt: phase A counts, phase B zeroed.  After
      // phase A it is already so (phase A sets neither an error nor
      // max_ulen, or the job has ended), and the decoder, which needs only
      // phase A's arrays, is launched first (or ran behind phase A's
      // read-back): it starts as soon as the host has seen phase A end, and
      // the host tables are staged behind it.
      DevStatus sb = stA;
      sb.err = 0;
      sb.site = 0;
      sb.max_ulen = 0;
      const bool decode_first = !resident && attempt == 1 && stA.err == 0 && stA.site == 0 && stA.max_ulen == 0;
      auto decode = [&] {
        if (!use_guess)
          launch_decode(T, d_files, d_blks, nb, d_ebase, K, vptr, vlen, c->st, stA.max_blk, stA.max_nr, stA.max_ent,
                        tune.decode_groups);
      };
      if (decode_first) decode();
      // the fused last merge round's crossing slots start empty (filled
      // here rather than by a launch of their own)
      const StageFill slots_fill = W <= 8 && ngp ? StageFill{events, (u64)ngp + 1, 0xffffffffu} : StageFill{nullptr, 0, 0};
      stage_in(c, d_tab, blob.data(), blob.size(), decode_first ? nullptr : &sb, slots_fill);
      // decode, check, merge, drop, gather, stops
      HIPCHK(scan_reserve(c->scans,
                          std::max<u64>(std::max<u64>(n, es.m_cap), std::max<u64>(es.nbo_cap, std::max(es.nf_cap, es.nw_cap))) + 1,
                          c->s));
      // the fused last merge round looks back over its merge tiles, the
      // table layout over the tables
      HIPCHK(scan_reserve_tiles(c->scans,
                                std::max<u64>(((u64)n + merge_items_per_tile(W) - 1) / merge_items_per_tile(W) + 1,
                                              (u64)es.nf_cap + 1),
                                c->s));
      if (resident) launch_gather_decoded(T, d_dsrc, nfiles, n, K, vptr, vlen);
      else if (!decode_first) decode();
      if (W <= 8) {
        // Every round but the last writes a permutation (the first reads the
        // children in place and checks they are sorted); the last round
        // selects and writes the kept stream itself.
        const u32* in = nullptr;
        u32* bufs[2] = {permA, permB};
        u32 nbuf = 0;
        u64 poff = 0;
        const u32 items = merge_items_per_tile(W);
        for (size_t ri = 0; ri + 1 < rounds.size(); ri++) {
          const auto& r = rounds[ri];
          const MergePair& last = r.back();
          const u32 tiles = last.tile0 + (last.b1 - last.a0 + items - 1) / items;
          launch_merge(T, d_pairs + poff, (u32)r.size(), tiles, in, bufs[nbuf], K, ri == 0, splits, c->st);
          in = bufs[nbuf];
          nbuf ^= 1;
          poff += r.size();
        }
        launch_merge_select(T, c->scans, d_pairs + (allp.size() - 1), fin, in, rounds.size() <= 1, K, vptr, vlen, dp,
                            EA.kept, gcnt, kpos, events, splits, c->st);
      } else {
        launch_check_sorted(T, K, n, d_cstart, (u32)cstart.size() - 1, permA, c->st);
        u32* in = permA;
        u32* outp = permB;
        u64 poff = 0;
        for (auto& r : rounds) {
          const MergePair& last = r.back();
          const u32 items = merge_items_per_tile(W);
          const u32 tiles = last.tile0 + (last.b1 - last.a0 + items - 1) / items;
          launch_merge(T, d_pairs + poff, (u32)r.size(), tiles, in, outp, K, false, splits, c->st);
          poff += r.size();
          std::swap(in, outp);
        }
        launch_drop(T, in, n, K, dp, keep, gcnt, events, c->st);
        scan_keep(T, c->scans, keep, kpos, n, dp, gcnt, events, ngp, c->st);
        launch_gather(T, in, n, keep, kpos, K, vptr, vlen, EA.kept);
      }
      if (ngp) launch_stops(T, gcnt, kpos, keep, n, dp, events, ngp, EA.kept, c->st);
      run_encode(c, es, EA, (u8*)out);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev1, c->s));
    // status, table metadata and boundary keys (at capacity) in one round trip
    const u64 o_meta = align_up(sizeof(DevStatus), 64);
    const u64 meta_bytes = (u64)es.nf_cap * sizeof(FileMeta);
    const u64 o_keys = align_up(o_meta + meta_bytes, 64);
    const u64 key_bytes = (u64)es.nf_cap * 2 * es.ep.kw;
    HIPCHK(c->pin.reserve(o_keys + key_bytes + 64));
    stage_out(c, StageSeg{c->st, 0, sizeof(DevStatus)}, StageSeg{EA.eb.meta, o_meta, meta_bytes},
              StageSeg{EA.eb.keys, o_keys, key_bytes});
    stage_sync(c);
    DevStatus stB;
    memcpy(&stB, c->pin.p, sizeof(DevStatus));
    float dms = 0;
    hipEventElapsedTime(&dms, c->ev0, c->ev1);
    stats.device_ms = dms;
    // An overflow (keys wider than the planes, a capacity guessed too low)
    // is retried first: what the kernels after it reported may come from the
    // truncated keys or clipped counts; a real fault shows again next time.
    const bool overflow = (stB.err & ERR_OVERFLOW) || stB.out_bytes > out_cap;
    if (status_code_of(stB) != LUDA_OK && (!overflow || attempt >= kMaxAttempts)) {
      c->give_out(out);
      throw Fail{status_code_of(stB), std::string("compaction rejected: ") + site_name(stB.site)};
    }
    if (overflow) {
      c->give_out(out);
      if (attempt >= kMaxAttempts) throw Fail{LUDA_NOT_SUPPORTED, std::string("capacity: ") + site_name(stB.site)};
      bool grew = false;
      if (stB.max_ulen > 8 * W) {
        W = (stB.max_ulen + 7) / 8;
        grew = true;
      }
      if (stB.out_bytes > out_cap) {
        out_cap = stB.out_bytes + (1 << 20);
        grew = true;
      }
      if (stB.need_files) {
        nf_min = (u64)stB.need_files + 16;
        grew = true;
      }
      if (stB.need_windows) {
        nw_min = (u64)stB.need_windows + 16;
        grew = true;
      }
      if (stB.need_blocks) {
        nbo_min = (u64)stB.need_blocks + 16;
        grew = true;
      }
      if (stB.cut_fallback && !generic_cuts) {
        generic_cuts = true;
        grew = true;
      }
      if (!grew) {
        const int code = status_code_of(stB);
        throw Fail{code != LUDA_OK ? code : LUDA_NOT_SUPPORTED, std::string("capacity: ") + site_name(stB.site)};
      }
      continue;
    }
    c->kw_hint = std::max<u32>(c->kw_hint, (stB.max_ulen + 7) / 8);
    stats.attempts = attempt;
    stats.decode_ahead = use_guess ? 1 : 0;
    stats.key_width = 8 * W;
    stats.entries_kept = stB.n_kept;
    stats.data_block_bytes = stB.data_bytes;
    fill_kernel_stats(c, &stats, W, rounds.size());
    luda_result* r = finish_result(c, es, EA, out, stB, job->flags, stats,
                                   reinterpret_cast<const FileMeta*>(c->pin.p + o_meta), c->pin.p + o_keys);
    if (job->flags & LUDA_JOB_KEEP_DECODED) {
      try {
        keep_decoded(c, EA, W, stB.n_kept, r);
      } catch (...) {
        free_result_now(c, r);
        throw;
      }
    }
    r->stats.total_ms = now_ms() - t_call;
    r->stats.end_clock_ms = now_ms();
    return r;
  }
}

// ---------------------------------------------------------------------------
// luda_verify_tables

void verify_tables(luda_ctx* c, const luda_file* tables, u32 n, u32 flags, u32* first_bad) {
  *first_bad = 0xffffffffu;
  if (n == 0) return;
  std::vector<FileDesc> fdesc(n);
  u64 nb_cap = 0;
  if (flags & LUDA_JOB_INPUTS_ON_DEVICE) {
    for (u32 i = 0; i < n; i++) fdesc[i] = {tables[i].data, tables[i].size};
  } else {
    u64 need = 0;
    for (u32 i = 0; i < n; i++) need += align_up(tables[i].size + 16, 256);
    HIPCHK(c->in_buf.reserve(need));
    u64 off = 0;
    for (u32 i = 0; i < n; i++) {
      HIPCHK(hipMemcpyAsync(c->in_buf.p + off, tables[i].data, tables[i].size, hipMemcpyHostToDevice, c->s));
      fdesc[i] = {c->in_buf.p + off, tables[i].size};
      off += align_up(tables[i].size + 16, 256);
    }
  }
  for (u32 i = 0; i < n; i++) nb_cap += tables[i].size / 41 + 1;
  if (nb_cap > 0xffffffffull / 2) throw Fail{LUDA_NOT_SUPPORTED, "tables too large"};
  FileDesc* d_files;
  FileInfo* d_info;
  u32 *d_bbase, *d_nent;
  BlockDesc* d_blks;
  auto carve = [&](Arena& a) {
    d_files = a.take<FileDesc>(n);
    d_info = a.take<FileInfo>(n);
    d_bbase = a.take<u32>(n + 1);
    d_blks = a.take<BlockDesc>(nb_cap);
    d_nent = a.take<u32>(nb_cap + 1);
  };
  {
    Arena dry;
    dry.dry = true;
    carve(dry);
    HIPCHK(c->a_buf.reserve(dry.used));
    Arena a;
    a.base = c->a_buf.p;
    a.cap = c->a_buf.cap;
    carve(a);
  }
  init_ctx_timer(c);
  KTimer& T = c->T;
  DevStatus st0;
  memset(&st0, 0, sizeof(st0));
  st0.bad_file = 0xffffffffu;
  HIPCHK(scan_reserve(c->scans, std::max<u64>(nb_cap, n + 1), c->s));
  files_in(c, d_files, fdesc.data(), n, st0, d_info, d_bbase);
  launch_index(T, d_files, d_info, d_bbase, n, (u32)nb_cap, d_blks, d_nent, c->st);
  launch_verify(T, d_files, n, d_blks, (u32)nb_cap, c->st);
  HIPCHK(hipGetLastError());
  stage_out(c, StageSeg{c->st, 0, sizeof(DevStatus)});
  stage_sync(c);
  DevStatus st;
  memcpy(&st, c->pin.p, sizeof(DevStatus));
  *first_bad = st.bad_file;
  if (status_code_of(st) != LUDA_OK)
    throw Fail{status_code_of(st), std::string("table rejected: ") + site_name(st.site)};
}

// ---------------------------------------------------------------------------
// luda_read_tables

luda_entries* read_tables(luda_ctx* c, const luda_file* tables, u32 nfiles, u32 flags) {
  luda_entries* r = (luda_entries*)calloc(1, sizeof(luda_entries));
  if (!r) throw std::bad_alloc();
  struct Owner {
    luda_entries* r;
    ~Owner() {
      if (r) luda_free_entries(r);
    }
  } own{r};
  r->num_tables = nfiles;
  r->table_first = (uint64_t*)calloc(nfiles + 1, sizeof(uint64_t));
  r->key_off = (uint64_t*)calloc(1, sizeof(uint64_t));
  r->val_off = (uint64_t*)calloc(1, sizeof(uint64_t));
  if (!r->table_first || !r->key_off || !r->val_off) throw std::bad_alloc();
  if (nfiles == 0) {
    own.r = nullptr;
    return r;
  }
  HIPCHK(c->pin.reserve(1024 + 4ull * (nfiles + 1) + sizeof(DevStatus)));
  init_ctx_timer(c);
  KTimer& T = c->T;
  std::vector<FileDesc> fdesc(nfiles);
  u64 bytes_in = 0, nb_cap = 0;
  for (u32 i = 0; i < nfiles; i++) {
    bytes_in += tables[i].size;
    nb_cap += tables[i].size / 41 + 1;
  }
  if (nb_cap > 0xffffffffull / 2) throw Fail{LUDA_NOT_SUPPORTED, "tables too large"};
  if (flags & LUDA_JOB_INPUTS_ON_DEVICE) {
    for (u32 i = 0; i < nfiles; i++) fdesc[i] = {tables[i].data, tables[i].size};
  } else {
    u64 need = 0;
    for (u32 i = 0; i < nfiles; i++) need += align_up(tables[i].size + 16, 256);
    HIPCHK(c->in_buf.reserve(need));
    u64 off = 0;
    for (u32 i = 0; i < nfiles; i++) {
      HIPCHK(hipMemcpyAsync(c->in_buf.p + off, tables[i].data, tables[i].size, hipMemcpyHostToDevice, c->s));
      fdesc[i] = {c->in_buf.p + off, tables[i].size};
      off += align_up(tables[i].size + 16, 256);
    }
  }
  // footers, index blocks and restart arrays: blocks and entries per table
  FileDesc* d_files;
  FileInfo* d_info;
  u32 *d_bbase, *d_nent, *d_ebase, *d_febase;
  BlockDesc* d_blks;
  auto carve_a = [&](Arena& a) {
    d_files = a.take<FileDesc>(nfiles);
    d_info = a.take<FileInfo>(nfiles);
    d_bbase = a.take<u32>(nfiles + 1);
    d_blks = a.take<BlockDesc>(nb_cap);
    d_nent = a.take<u32>(nb_cap + 1);
    d_ebase = a.take<u32>(nb_cap + 1);
    d_febase = a.take<u32>(nfiles + 1);
  };
  {
    Arena dry;
    dry.dry = true;
    carve_a(dry);
    HIPCHK(c->a_buf.reserve(dry.used));
    Arena a;
    a.base = c->a_buf.p;
    a.cap = c->a_buf.cap;
    carve_a(a);
  }
  DevStatus st0;
  memset(&st0, 0, sizeof(st0));
  st0.bad_file = 0xffffffffu;
  HIPCHK(scan_reserve_tiles(c->scans, index_scan_tiles((u32)nb_cap), c->s));
  files_in(c, d_files, fdesc.data(), nfiles, st0, d_info, d_bbase);
  launch_index_entries(T, c->scans, d_files, d_info, d_bbase, nfiles, (u32)nb_cap, d_blks, d_nent, d_ebase, d_febase,
                       c->st);
  if (flags & LUDA_READ_VERIFY) launch_verify(T, d_files, nfiles, d_blks, (u32)nb_cap, c->st);
  stage_out(c, StageSeg{c->st, 0, sizeof(DevStatus)}, StageSeg{d_febase, 1024, (nfiles + 1) * 4ull});
  stage_sync(c);
  DevStatus stA;
  memcpy(&stA, c->pin.p, sizeof(DevStatus));
  std::vector<u32> febase(nfiles + 1);
  memcpy(febase.data(), c->pin.p + 1024, (nfiles + 1) * 4);
  if (status_code_of(stA) != LUDA_OK)
    throw Fail{status_code_of(stA), std::string("table rejected: ") + site_name(stA.site)};
  const u32 nb = stA.n_blocks_in;
  const u64 n = stA.n_entries;
  if (nb > nb_cap) throw Fail{LUDA_CORRUPTION, "more data blocks than the tables can hold"};
  for (u32 i = 0; i <= nfiles; i++) r->table_first[i] = febase[i];
  // decode, then the entries' bytes to plain arrays
  u32 W = std::max<u32>(1, (std::max<u32>(stA.max_rulen, 8 * c->kw_hint) + 7) / 8);
  for (u32 attempt = 1;; attempt++) {
    if (W > kMaxKeyWords) throw Fail{LUDA_NOT_SUPPORTED, "user keys longer than the device key planes"};
    Planes K;
    u64 *vptr, *key_off, *val_off;
    u32 *vlen, *ksz;
    u8 *keys, *vals;
    auto carve_b = [&](Arena& a) {
      K.P = a.take<u64>((u64)W * std::max<u64>(n, 1));
      K.klen = a.take<u32>(n + 1);
      K.tag = a.take<u64>(n + 1);
      K.stride = std::max<u64>(n, 1);
      K.W = W;
      vptr = a.take<u64>(n + 1);
      vlen = a.take<u32>(n + 1);
      ksz = a.take<u32>(n + 1);
      key_off = a.take<u64>(n + 1);
      val_off = a.take<u64>(n + 1);
      keys = a.take<u8>(n * (8ull * W + 8) + 16);
      vals = a.take<u8>(bytes_in + 16);
    };
    {
      Arena dry;
      dry.dry = true;
      carve_b(dry);
      if (dry.used > c->budget) throw Fail{LUDA_NOT_SUPPORTED, "tables exceed the HBM budget"};
      HIPCHK(c->b_buf.reserve(dry.used));
      Arena a;
      a.base = c->b_buf.p;
      a.cap = c->b_buf.cap;
      carve_b(a);
    }
    DevStatus sb = stA;
    sb.err = 0;
    sb.site = 0;
    sb.max_ulen = 0;
    stage_in(c, nullptr, nullptr, 0, &sb);
    HIPCHK(scan_reserve(c->scans, n + 2, c->s));
    launch_decode(T, d_files, d_blks, nb, d_ebase, K, vptr, vlen, c->st, stA.max_blk, stA.max_nr, stA.max_ent);
    launch_entry_out_sizes(T, K, n, ksz, vlen);
    scan_u32_to_u64(T, c->scans, ksz, key_off, nullptr, (u32)n + 1, (u32)n + 1, nullptr, c->st);
    scan_u32_to_u64(T, c->scans, vlen, val_off, nullptr, (u32)n + 1, (u32)n + 1, nullptr, c->st);
    launch_emit(T, K, vptr, vlen, n, key_off, val_off, keys, vals);
    HIPCHK(hipGetLastError());
    constexpr u64 o_kb = align_up_c(sizeof(DevStatus), 64), o_vb = o_kb + 64;
    stage_out(c, StageSeg{c->st, 0, sizeof(DevStatus)}, StageSeg{key_off + n, o_kb, 8}, StageSeg{val_off + n, o_vb, 8});
    stage_sync(c);
    DevStatus stB;
    memcpy(&stB, c->pin.p, sizeof(DevStatus));
    if (stB.err & ERR_OVERFLOW) {
      if (attempt >= kMaxAttempts || stB.max_ulen <= 8 * W)
        throw Fail{LUDA_NOT_SUPPORTED, std::string("capacity: ") + site_name(stB.site)};
      W = (stB.max_ulen + 7) / 8;
      continue;
    }
    if (status_code_of(stB) != LUDA_OK)
      throw Fail{status_code_of(stB), std::string("table rejected: ") + site_name(stB.site)};
    c->kw_hint = std::max<u32>(c->kw_hint, (stB.max_ulen + 7) / 8);
    u64 kb = 0, vb = 0;
    memcpy(&kb, c->pin.p + o_kb, 8);
    memcpy(&vb, c->pin.p + o_vb, 8);
    free(r->key_off);
    free(r->val_off);
    r->key_off = (uint64_t*)malloc((n + 1) * 8);
    r->val_off = (uint64_t*)malloc((n + 1) * 8);
    r->keys = (uint8_t*)malloc(kb ? kb : 1);
    r->vals = (uint8_t*)malloc(vb ? vb : 1);
    if (!r->key_off || !r->val_off || !r->keys || !r->vals) throw std::bad_alloc();
    HIPCHK(hipMemcpyAsync(r->key_off, key_off, (n + 1) * 8, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(r->val_off, val_off, (n + 1) * 8, hipMemcpyDeviceToHost, c->s));
    if (kb) HIPCHK(hipMemcpyAsync(r->keys, keys, kb, hipMemcpyDeviceToHost, c->s));
    if (vb) HIPCHK(hipMemcpyAsync(r->vals, vals, vb, hipMemcpyDeviceToHost, c->s));
    stream_sync(c);
    r->n = n;
    own.r = nullptr;
    return r;
  }
}

// ---------------------------------------------------------------------------
// luda_verify_tables_noalloc: luda_verify_tables with its working memory in
// the code object.  Every device the library runs on holds its own copy of
// `scratch` until the process ends; host code reaches it only through the
// device address hipGetSymbolAddress returned for that device (noalloc_setup,
// called once per device by luda_init).  Layout: the look-back words of the
// call's one prefix sum, then the table descriptors, per-table index
// handles and block counts, and the data block descriptors.  Calls that
// overlap on one device share it and race.

}  // namespace

constexpr u64 kNoAllocScratch = 72ull << 20;
constexpr u64 kNoAllocMaxBytes = 128ull << 20;  // sum of table sizes per call
constexpr u32 kNoAllocMaxTables = 4096;
constexpr u64 kNoAllocScanTiles = 4096;         // look-back words and the tile ticket: scans of up to 16 Mi - 4 Ki items
static __device__ __align__(256) u8 scratch[kNoAllocScratch];

namespace {

constexpr int kMaxDevices = 64;
struct NoAllocDev {
  u8* ptr = nullptr;  // this device's copy of `scratch`
  ScanState scans;    // look-back words at ptr, epochs advanced under mu
};
std::mutex g_noalloc_mu;
NoAllocDev g_noalloc[kMaxDevices];

// With `device` current: its copy's address, and the look-back words zeroed
// (on stream s, which the caller synchronises).
void noalloc_setup(int device, hipStream_t s) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  NoAllocDev& d = g_noalloc[device];
  if (d.ptr) return;
  void* p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw Fail{LUDA_IO_ERROR, "no device address for the scratch array"};
  HIPCHK(scan_attach(d.scans, reinterpret_cast<u64*>(p), kNoAllocScanTiles, s));
  HIPCHK(hipStreamSynchronize(s));
  d.ptr = static_cast<u8*>(p);
}

void verify_tables_noalloc(luda_ctx* c, const luda_file* tables, u32 n, u32 flags, u32* first_bad) {
  *first_bad = 0xffffffffu;
  if (!(flags & LUDA_JOB_INPUTS_ON_DEVICE))
    throw Fail{LUDA_INVALID_ARGUMENT, "luda_verify_tables_noalloc takes device-resident tables only"};
  if (n == 0) return;
  if (n > kNoAllocMaxTables) throw Fail{LUDA_NOT_SUPPORTED, "more than 4096 tables in one call"};
  u64 bytes = 0, nb_cap = 0;
  for (u32 i = 0; i < n; i++) {
    bytes += tables[i].size;
    nb_cap += tables[i].size / 41 + 1;
  }
  if (bytes > kNoAllocMaxBytes) throw Fail{LUDA_NOT_SUPPORTED, "more than 128 MiB of tables in one call"};
  if (c->device < 0 || c->device >= kMaxDevices || !g_noalloc[c->device].ptr)
    throw Fail{LUDA_IO_ERROR, "scratch array not set up on this device"};
  NoAllocDev& D = g_noalloc[c->device];
  FileDesc* d_files;
  FileInfo* d_info;
  u32 *d_bbase, *d_nent;
  BlockDesc* d_blks;
  Arena a;
  a.base = D.ptr + align_up(kNoAllocScanTiles * sizeof(u64), kAlign);
  a.cap = kNoAllocScratch - align_up(kNoAllocScanTiles * sizeof(u64), kAlign);
  d_files = a.take<FileDesc>(n);
  d_info = a.take<FileInfo>(n);
  d_bbase = a.take<u32>(n + 1);
  d_blks = a.take<BlockDesc>(nb_cap);
  d_nent = a.take<u32>(nb_cap + 1);
  if (a.used > a.cap) throw Fail{LUDA_NOT_SUPPORTED, "tables need more block slots than the scratch array holds"};
  // descriptors and the cleared status through the pinned staging buffer
  // that luda_init reserved, written in place (no host array either)
  if (c->pin2_busy) stream_sync(c);
  const u64 o_st = align_up((u64)n * sizeof(FileDesc), 64);
  if (o_st + sizeof(DevStatus) > c->pin2.cap) throw Fail{LUDA_IO_ERROR, "staging buffer too small"};
  FileDesc* hd = reinterpret_cast<FileDesc*>(c->pin2.p);
  for (u32 i = 0; i < n; i++) hd[i] = FileDesc{tables[i].data, tables[i].size};
  DevStatus* hst = reinterpret_cast<DevStatus*>(c->pin2.p + o_st);
  memset(hst, 0, sizeof(DevStatus));
  hst->bad_file = 0xffffffffu;
  KTimer T;  // no per-kernel timing: it would create events
  T.s = c->s;
  launch_files_in(T, reinterpret_cast<const FileDesc*>(c->pin2.d), reinterpret_cast<const DevStatus*>(c->pin2.d + o_st),
                  d_files, n, d_info, d_bbase, c->st);
  c->pin2_busy = true;
  launch_index(T, d_files, d_info, d_bbase, n, (u32)nb_cap, d_blks, d_nent, c->st);
  launch_verify(T, d_files, n, d_blks, (u32)nb_cap, c->st);
  HIPCHK(hipGetLastError());
  stage_out(c, StageSeg{c->st, 0, sizeof(DevStatus)});
  stage_sync(c);
  DevStatus st;
  memcpy(&st, c->pin.p, sizeof(DevStatus));
  *first_bad = st.bad_file;
  if (status_code_of(st) != LUDA_OK) throw Fail{status_code_of(st), "table rejected"};
}

// ---------------------------------------------------------------------------
// luda_build_tables

luda_result* build_tables(luda_ctx* c, const luda_options* opts, u64 n, const u8* keys, const u64* key_off,
                          const u8* vals, const u64* val_off, u32 flags) {
  const double t_call = now_ms();
  luda_stats stats;
  memset(&stats, 0, sizeof(stats));
  validate_options(*opts);
  if (n >= 0xffffffffull) throw Fail{LUDA_NOT_SUPPORTED, "too many entries"};
  u64 maxu = 0;
  for (u64 i = 0; i < n; i++) {
    const u64 kl = key_off[i + 1] - key_off[i];
    if (kl < 8) throw Fail{LUDA_NOT_SUPPORTED, "internal key shorter than 8 bytes"};
    maxu = std::max<u64>(maxu, kl - 8);
  }
  const u32 W = (u32)std::max<u64>(1, (maxu + 7) / 8);
  if (W > kMaxKeyWords) throw Fail{LUDA_NOT_SUPPORTED, "user keys longer than the device key planes"};
  HIPCHK(c->pin.reserve(1 << 16));
  init_ctx_timer(c);
  const u64 kbytes = key_off[n], vbytes = val_off[n];
  stats.entries_in = n;
  if (n == 0) {
    luda_result* r = (luda_result*)calloc(1, sizeof(luda_result));
    r->outputs = (luda_output*)calloc(1, sizeof(luda_output));
    r->keys = (uint8_t*)malloc(1);
    r->stats = stats;
    r->stats.attempts = 1;
    r->stats.end_clock_ms = now_ms();
    return r;
  }
  // inputs: keys, offsets and values uploaded once
  u8 *d_keys, *d_vals;
  u64 *d_koff, *d_voff;
  {
    Arena dry;
    dry.dry = true;
    dry.take<u8>(kbytes + 16);
    dry.take<u8>(vbytes + 16);
    dry.take<u64>(n + 1);
    dry.take<u64>(n + 1);
    HIPCHK(c->in_buf.reserve(dry.used));
    Arena a;
    a.base = c->in_buf.p;
    a.cap = c->in_buf.cap;
    d_keys = a.take<u8>(kbytes + 16);
    d_vals = a.take<u8>(vbytes + 16);
    d_koff = a.take<u64>(n + 1);
    d_voff = a.take<u64>(n + 1);
  }
  const double t = now_ms();
  HIPCHK(hipMemcpyAsync(d_keys, keys, kbytes, hipMemcpyHostToDevice, c->s));
  HIPCHK(hipMemcpyAsync(d_vals, vals, vbytes, hipMemcpyHostToDevice, c->s));
  HIPCHK(hipMemcpyAsync(d_koff, key_off, (n + 1) * 8, hipMemcpyHostToDevice, c->s));
  HIPCHK(hipMemcpyAsync(d_voff, val_off, (n + 1) * 8, hipMemcpyHostToDevice, c->s));
  stream_sync(c);
  stats.h2d_ms = now_ms() - t;
  stats.bytes_read = kbytes + vbytes;
  HIPCHK(hipEventRecord(c->ev0, c->s));
  u64 out_cap = kbytes + vbytes + 16 * n + (1 << 20);
  u64 nf_min = 0, nw_min = 0, nbo_min = 0;
  bool generic_cuts = start_generic_cuts(*opts, n, kbytes + vbytes);
  for (u32 attempt = 1;; attempt++) {
    EncodeSetup es = plan_encode(*opts, W, n, out_cap, 0, nf_min, nw_min, nbo_min);
    es.wb_groups = pick_tuning(c, kbytes + vbytes, 8 * W).write_blocks_groups;
    es.generic_cuts = generic_cuts;
    EncodeArrays EA;
    {
      Arena dry;
      dry.dry = true;
      carve_encode(dry, es, W, true, (flags & LUDA_JOB_KEEP_DECODED) != 0);
      if (dry.used > c->budget) throw Fail{LUDA_NOT_SUPPORTED, "job exceeds the HBM budget"};
      HIPCHK(c->b_buf.reserve(dry.used));
      Arena a;
      a.base = c->b_buf.p;
      a.cap = c->b_buf.cap;
      EA = carve_encode(a, es, W, true, (flags & LUDA_JOB_KEEP_DECODED) != 0);
    }
    void* out = c->take_out(out_cap);
    if (!out) throw Fail{LUDA_IO_ERROR, "output allocation failed"};
    DevStatus s0;
    memset(&s0, 0, sizeof(s0));
    s0.n_entries = (u32)n;
    s0.n_kept = (u32)n;
    stage_in(c, nullptr, nullptr, 0, &s0);
    launch_ingest(c->T, n, d_keys, d_koff, d_vals, d_voff, EA.kept, c->st);
    run_encode(c, es, EA, (u8*)out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev1, c->s));
    // status, table metadata and boundary keys (at capacity) in one round trip
    const u64 o_meta = align_up(sizeof(DevStatus), 64);
    const u64 meta_bytes = (u64)es.nf_cap * sizeof(FileMeta);
    const u64 o_keys = align_up(o_meta + meta_bytes, 64);
    const u64 key_bytes = (u64)es.nf_cap * 2 * es.ep.kw;
    HIPCHK(c->pin.reserve(o_keys + key_bytes + 64));
    stage_out(c, StageSeg{c->st, 0, sizeof(DevStatus)}, StageSeg{EA.eb.meta, o_meta, meta_bytes},
              StageSeg{EA.eb.keys, o_keys, key_bytes});
    stage_sync(c);
    DevStatus stB;
    memcpy(&stB, c->pin.p, sizeof(DevStatus));
    float dms = 0;
    hipEventElapsedTime(&dms, c->ev0, c->ev1);
    stats.device_ms = dms;
    const bool overflow = (stB.err & ERR_OVERFLOW) || stB.out_bytes > out_cap;
    if (status_code_of(stB) != LUDA_OK && (!overflow || attempt >= kMaxAttempts)) {
      c->give_out(out);
      throw Fail{status_code_of(stB), std::string("build rejected: ") + site_name(stB.site)};
    }
    if (overflow) {
      c->give_out(out);
      if (attempt >= kMaxAttempts) throw Fail{LUDA_NOT_SUPPORTED, std::string("capacity: ") + site_name(stB.site)};
      if (stB.out_bytes > out_cap) out_cap = stB.out_bytes + (1 << 20);
      else out_cap *= 2;
      if (stB.need_files) nf_min = (u64)stB.need_files + 16;
      if (stB.need_windows) nw_min = (u64)stB.need_windows + 16;
      if (stB.need_blocks) nbo_min = (u64)stB.need_blocks + 16;
      if (stB.cut_fallback) generic_cuts = true;
      continue;
    }
    stats.attempts = attempt;
    stats.key_width = 8 * W;
    stats.entries_kept = stB.n_kept;
    stats.data_block_bytes = stB.data_bytes;
    fill_kernel_stats(c, &stats, W, 0);
    luda_result* r = finish_result(c, es, EA, out, stB, flags, stats,
                                   reinterpret_cast<const FileMeta*>(c->pin.p + o_meta), c->pin.p + o_keys);
    if (flags & LUDA_JOB_KEEP_DECODED) {
      try {
        keep_decoded(c, EA, W, stB.n_kept, r);
      } catch (...) {
        free_result_now(c, r);
        throw;
      }
    }
    r->stats.total_ms = now_ms() - t_call;
    r->stats.end_clock_ms = now_ms();
    return r;
  }
}

// LUDA_FAULT=<site>:<kind> makes the call at <site> ("init", "compact",
// "build", "verify") throw before it does any work: kind "bad_alloc" throws
// std::bad_alloc, "runtime" a std::runtime_error, anything else an int.  It
// lets a test check that no exception crosses the C-ABI.
void fault_point(const char* site) {
  const char* e = getenv("LUDA_FAULT");
  if (!e || !*e) return;
  const char* colon = strchr(e, ':');
  const size_t n = colon ? (size_t)(colon - e) : strlen(e);
  if (strlen(site) != n || strncmp(e, site, n) != 0) return;
  const char* kind = colon ? colon + 1 : "";
  if (!strcmp(kind, "bad_alloc")) throw std::bad_alloc();
  if (!strcmp(kind, "runtime")) throw std::runtime_error("injected fault");
  throw 42;
}

// Every entry point runs its body through here, so that no C++ exception
// crosses the C-ABI (a caller's background thread would end in
// std::terminate): the engine's own failures keep their status; running out
// of host memory, or any other exception, is an IO_ERROR, which sends the
// caller to its CPU path like every other non-OK status.
template <typename F>
int guarded(luda_ctx* c, char* err, size_t errlen, F fn) {
  int code = LUDA_OK;
  try {
    fn();
    return LUDA_OK;
  } catch (const Fail& f) {
    if (err && errlen) snprintf(err, errlen, "%s", f.msg.c_str());
    code = f.code;
  } catch (const std::bad_alloc&) {
    if (err && errlen) snprintf(err, errlen, "host memory exhausted");
    code = LUDA_IO_ERROR;
  } catch (const std::exception& x) {
    if (err && errlen) snprintf(err, errlen, "internal error: %s", x.what());
    code = LUDA_IO_ERROR;
  } catch (...) {
    if (err && errlen) snprintf(err, errlen, "internal error");
    code = LUDA_IO_ERROR;
  }
  // a failed call may leave work queued; drain it so the next call starts clean
  if (c && c->s) {
    hipStreamSynchronize(c->s);
    c->pin2_busy = false;
  }
  return code;
}

// The worker slots of luda_submit.  Each slot is a context of its own
// (stream, working memory, output buffers) driven by one host thread; the
// queue admits a job when a slot is idle and the estimated device memory of
// the running jobs plus this one fits the budget (or nothing runs).
// A flush queued by luda_submit_build: the arguments of luda_build_tables.
struct BuildReq {
  luda_options opts;
  u64 n;
  const u8* keys;
  const u64* key_off;
  const u8* vals;
  const u64* val_off;
  u32 flags;
};
struct QueuedJob {
  const luda_job* job;  // nullptr for a flush
  BuildReq build;
  luda_done_fn done;
  void* user;
  u64 est;  // estimated device bytes while it runs
};
struct DoneJob {
  luda_done_fn done;
  void* user;
  int status;
  luda_result* result;
  std::string err;
  luda_ctx* owner;
};

u64 job_footprint(const luda_job* j) {
  u64 in = 0;
  for (int w = 0; w < 2; w++)
    for (u32 i = 0; i < j->num_inputs[w]; i++) in += j->inputs[w][i].size;
  return in * 9 / 2 + (64ull << 20);
}

u64 build_footprint(const BuildReq& b) {
  const u64 in = b.n ? b.key_off[b.n] + b.val_off[b.n] : 0;
  return in * 9 / 2 + (64ull << 20);
}

}  // namespace

struct JobPool {
  std::vector<luda_ctx*> ctxs;
  std::vector<std::thread> threads;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<QueuedJob> queue;
  std::deque<DoneJob> done;
  u64 budget = 0, inflight = 0;
  u32 running = 0, submitted = 0, delivered = 0;
  bool stop = false;
};

namespace {

int ctx_open(int device, u64 budget, luda_ctx** out, char* err, size_t errlen);
void ctx_close(luda_ctx* c);
int ctx_compact(luda_ctx* c, const luda_job* job, luda_result** out, char* err, size_t errlen);
int ctx_build(luda_ctx* c, const luda_options* opts, u64 n, const u8* keys, const u64* key_off, const u8* vals,
              const u64* val_off, u32 flags, luda_result** out, char* err, size_t errlen);

void pool_worker(JobPool* P, luda_ctx* w) {
  hipSetDevice(w->device);
  std::unique_lock<std::mutex> lk(P->mu);
  for (;;) {
    P->cv_work.wait(lk, [&] {
      return P->stop || (!P->queue.empty() && (P->running == 0 || P->inflight + P->queue.front().est <= P->budget));
    });
    if (P->stop && P->queue.empty()) return;
    QueuedJob q = P->queue.front();
    P->queue.pop_front();
    P->running++;
    P->inflight += q.est;
    lk.unlock();
    luda_result* r = nullptr;
    char err[256] = {0};
    const BuildReq& b = q.build;
    const int rc = q.job ? ctx_compact(w, q.job, &r, err, sizeof(err))
                         : ctx_build(w, &b.opts, b.n, b.keys, b.key_off, b.vals, b.val_off, b.flags, &r, err,
                                     sizeof(err));
    lk.lock();
    P->running--;
    P->inflight -= q.est;
    P->done.push_back(DoneJob{q.done, q.user, rc, r, err, w});
    P->cv_done.notify_all();
    P->cv_work.notify_all();
  }
}

JobPool* pool_of(luda_ctx* c) {
  if (c->pool) return c->pool;
  JobPool* P = new JobPool;
  P->budget = c->budget;
  for (u32 i = 0; i < c->slots; i++) {
    luda_ctx* w = nullptr;
    char err[256];
    if (ctx_open(c->device, c->budget, &w, err, sizeof(err)) != LUDA_OK) break;
    w->family = c;
    w->scans.tickets = true;  // slots run scans side by side (scan.h, take_ticket)
    w->rules = c->rules;
    w->use_heap(c->heap);
    P->ctxs.push_back(w);
  }
  if (P->ctxs.empty()) {
    delete P;
    return nullptr;
  }
  for (luda_ctx* w : P->ctxs) P->threads.emplace_back(pool_worker, P, w);
  c->pool = P;
  return P;
}

void pool_close(JobPool* P) {
  {
    std::lock_guard<std::mutex> lk(P->mu);
    P->stop = true;
  }
  P->cv_work.notify_all();
  for (auto& t : P->threads) t.join();
  for (auto& d : P->done)
    if (d.result) {
      // undelivered results: return their buffers before the slots go
      forget_decoded(d.result);
      if (d.result->device_data) d.owner->give_out(d.result->device_data);
      if (d.result->data) hipHostFree(d.result->data);
      free(d.result->outputs);
      free(d.result->keys);
      free(d.result);
    }
  for (luda_ctx* w : P->ctxs) ctx_close(w);
  delete P;
}

int pool_deliver(JobPool* P, bool wait_all) {
  int n = 0;
  for (;;) {
    DoneJob d;
    {
      std::unique_lock<std::mutex> lk(P->mu);
      if (wait_all) P->cv_done.wait(lk, [&] { return !P->done.empty() || P->delivered + P->done.size() == P->submitted; });
      if (P->done.empty()) return n;
      d = std::move(P->done.front());
      P->done.pop_front();
      P->delivered++;
    }
    if (d.done) d.done(d.user, d.status, d.result, d.err.c_str());
    n++;
  }
}

}  // namespace

extern "C" {

const char* luda_version(void) { return "luda-mi355x 0.1 (gfx950)"; }

int luda_init(int device, uint64_t hbm_budget_bytes, luda_ctx** out, char* err, size_t errlen) {
  return ctx_open(device, hbm_budget_bytes, out, err, errlen);
}

void luda_shutdown(luda_ctx* c) {
  if (!c) return;
  if (c->pool) pool_close(c->pool);
  c->pool = nullptr;
  ctx_close(c);
}

int luda_set_slots(luda_ctx* c, uint32_t slots) {
  if (!c || slots < 1 || slots > 16) return LUDA_INVALID_ARGUMENT;
  if (c->pool) return LUDA_INVALID_ARGUMENT;
  c->slots = slots;
  return LUDA_OK;
}

int luda_submit(luda_ctx* c, const luda_job* job, luda_done_fn done, void* user) {
  if (!c || !job) return LUDA_INVALID_ARGUMENT;
  return guarded(nullptr, nullptr, 0, [&] {
    JobPool* P = pool_of(c);
    if (!P) throw Fail{LUDA_IO_ERROR, "no worker slot"};
    {
      std::lock_guard<std::mutex> lk(P->mu);
      P->queue.push_back(QueuedJob{job, BuildReq{}, done, user, job_footprint(job)});
      P->submitted++;
    }
    P->cv_work.notify_one();
  });
}

int luda_submit_build(luda_ctx* c, const luda_options* opts, uint64_t n, const uint8_t* keys,
                      const uint64_t* key_off, const uint8_t* vals, const uint64_t* val_off, uint32_t flags,
                      luda_done_fn done, void* user) {
  if (!c || !opts || (n && (!keys || !key_off || !vals || !val_off))) return LUDA_INVALID_ARGUMENT;
  return guarded(nullptr, nullptr, 0, [&] {
    JobPool* P = pool_of(c);
    if (!P) throw Fail{LUDA_IO_ERROR, "no worker slot"};
    BuildReq b{*opts, n, keys, key_off, vals, val_off, flags};
    {
      std::lock_guard<std::mutex> lk(P->mu);
      P->queue.push_back(QueuedJob{nullptr, b, done, user, build_footprint(b)});
      P->submitted++;
    }
    P->cv_work.notify_one();
  });
}

int luda_set_tuning(luda_ctx* c, const luda_tuning_rule* rules, uint32_t n) {
  if (!c || (n && !rules)) return LUDA_INVALID_ARGUMENT;
  return guarded(nullptr, nullptr, 0, [&] {
    std::vector<luda_tuning_rule> r(rules, rules + n);
    {
      std::lock_guard<std::mutex> lk(c->mu);
      c->rules = r;
    }
    if (c->pool)
      for (luda_ctx* w : c->pool->ctxs) {
        std::lock_guard<std::mutex> lk(w->mu);
        w->rules = r;
      }
  });
}

int luda_poll(luda_ctx* c) {
  if (!c || !c->pool) return 0;
  return pool_deliver(c->pool, false);
}

int luda_drain(luda_ctx* c) {
  if (!c || !c->pool) return 0;
  return pool_deliver(c->pool, true);
}

}  // extern "C"

namespace {

int ctx_open(int device, uint64_t hbm_budget_bytes, luda_ctx** out, char* err, size_t errlen) {
  if (!out) return LUDA_INVALID_ARGUMENT;
  *out = nullptr;
  luda_ctx* c = new (std::nothrow) luda_ctx;
  if (!c) {
    if (err && errlen) snprintf(err, errlen, "host memory exhausted");
    return LUDA_IO_ERROR;
  }
  c->device = device;
  c->family = c;
  c->budget = hbm_budget_bytes ? hbm_budget_bytes : (64ull << 30);
  int rc = guarded(c, err, errlen, [&] {
    fault_point("init");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
      throw Fail{LUDA_NOT_SUPPORTED, "no HIP device " + std::to_string(device)};
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    c->use_heap(std::make_shared<DevHeap>());
    for (DevBuf* b : {&c->in_buf, &c->a_buf, &c->b_buf}) b->s = c->s;
    HIPCHK(hipMalloc(&c->st, sizeof(DevStatus)));
    HIPCHK(hipEventCreate(&c->ev0));
    HIPCHK(hipEventCreate(&c->ev1));
    HIPCHK(c->pin.reserve(1 << 20));
    HIPCHK(c->pin2.reserve(1 << 20));
    HIPCHK(hipHostMalloc(&c->done_h, 64, hipHostMallocCoherent));
    *c->done_h = 0;
    HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->done_d), c->done_h, 0));
    HIPCHK(hipMalloc(&c->done_ctr, 64));
    HIPCHK(hipMemset(c->done_ctr, 0, 64));
    c->T.s = c->s;
    noalloc_setup(device, c->s);
  });
  if (rc != LUDA_OK) {
    ctx_close(c);
    return rc;
  }
  *out = c;
  return LUDA_OK;
}

void ctx_close(luda_ctx* c) {
  if (!c) return;
  if (c->s) hipStreamSynchronize(c->s);
  if (c->io) io_close(c->io);
  c->io = nullptr;
  c->in_buf.release();
  c->a_buf.release();
  c->b_buf.release();
  scan_release(c->scans);
  c->pin.release();
  c->pin2.release();
  if (c->st) hipFree(c->st);
  if (c->done_ctr) hipFree(c->done_ctr);
  if (c->done_h) hipHostFree(c->done_h);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->s) hipStreamDestroy(c->s);
  delete c;
}

int ctx_compact(luda_ctx* c, const luda_job* job, luda_result** out, char* err, size_t errlen) {
  if (!c || !job || !out) return LUDA_INVALID_ARGUMENT;
  *out = nullptr;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  return guarded(c, err, errlen, [&] {
    fault_point("compact");
    *out = compact(c, job);
  });
}

int ctx_build(luda_ctx* c, const luda_options* opts, u64 n, const u8* keys, const u64* key_off, const u8* vals,
              const u64* val_off, u32 flags, luda_result** out, char* err, size_t errlen) {
  if (!c || !opts || !out || (n && (!keys || !key_off || !vals || !val_off))) return LUDA_INVALID_ARGUMENT;
  *out = nullptr;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  return guarded(c, err, errlen, [&] {
    fault_point("build");
    *out = build_tables(c, opts, n, keys, key_off, vals, val_off, flags);
  });
}

}  // namespace

extern "C" {

int luda_compact(luda_ctx* c, const luda_job* job, luda_result** out, char* err, size_t errlen) {
  return ctx_compact(c, job, out, err, errlen);
}

int luda_verify_tables(luda_ctx* c, const luda_file* tables, uint32_t n, uint32_t flags, uint32_t* first_bad,
                       char* err, size_t errlen) {
  if (!c || !first_bad || (n && !tables)) return LUDA_INVALID_ARGUMENT;
  *first_bad = 0xffffffffu;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  return guarded(c, err, errlen, [&] {
    fault_point("verify");
    verify_tables(c, tables, n, flags, first_bad);
  });
}

int luda_read_tables(luda_ctx* c, const luda_file* tables, uint32_t n, uint32_t flags, luda_entries** out, char* err,
                     size_t errlen) {
  if (!c || !out || (n && !tables)) return LUDA_INVALID_ARGUMENT;
  *out = nullptr;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  return guarded(c, err, errlen, [&] {
    fault_point("read");
    *out = read_tables(c, tables, n, flags);
  });
}

void luda_free_entries(luda_entries* e) {
  if (!e) return;
  free(e->keys);
  free(e->key_off);
  free(e->vals);
  free(e->val_off);
  free(e->table_first);
  free(e);
}

int luda_verify_tables_noalloc(luda_ctx* c, const luda_file* tables, uint32_t n, uint32_t flags, uint32_t* first_bad,
                               char* err, size_t errlen) {
  if (!c || !first_bad || (n && !tables)) return LUDA_INVALID_ARGUMENT;
  *first_bad = 0xffffffffu;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  return guarded(c, err, errlen, [&] {
    fault_point("verify");
    verify_tables_noalloc(c, tables, n, flags, first_bad);
  });
}

int luda_build_tables(luda_ctx* c, const luda_options* opts, uint64_t n, const uint8_t* keys, const uint64_t* key_off,
                      const uint8_t* vals, const uint64_t* val_off, uint32_t flags, luda_result** out, char* err,
                      size_t errlen) {
  return ctx_build(c, opts, n, keys, key_off, vals, val_off, flags, out, err, errlen);
}

void luda_free_result(luda_ctx* c, luda_result* r) {
  if (!r) return;
  forget_decoded(r);
  // the family's heap (the context and its worker slots share it) takes
  // the tables back; its own lock, so this does not wait for running jobs
  if (c && r->device_data) c->give_out(r->device_data);
  if (r->data) hipHostFree(r->data);
  free(r->outputs);
  free(r->keys);
  free(r);
}

int luda_device_alloc(luda_ctx* c, uint64_t bytes, void** dptr) {
  if (!c || !dptr) return LUDA_INVALID_ARGUMENT;
  hipSetDevice(c->device);
  return hipMalloc(dptr, bytes ? bytes : 1) == hipSuccess ? LUDA_OK : LUDA_IO_ERROR;
}
int luda_device_free(luda_ctx* c, void* dptr) {
  if (!c) return LUDA_INVALID_ARGUMENT;
  hipSetDevice(c->device);
  return hipFree(dptr) == hipSuccess ? LUDA_OK : LUDA_IO_ERROR;
}
int luda_memcpy_h2d(luda_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c) return LUDA_INVALID_ARGUMENT;
  hipSetDevice(c->device);
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? LUDA_OK : LUDA_IO_ERROR;
}
int luda_memcpy_d2h(luda_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c) return LUDA_INVALID_ARGUMENT;
  hipSetDevice(c->device);
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? LUDA_OK : LUDA_IO_ERROR;
}

}  // extern "C"

// io.hip — file I/O around a compaction and the install record (host code).
//
// luda_read_files: table files -> HBM.  The reference reads every input file
// whole with fread before decoding (cuda/cuda_common.h:302-311).  Here each
// file is read with pread in 8 MiB chunks into two pinned buffers that
// alternate, and each chunk is copied to HBM with hipMemcpyAsync on the
// context's I/O stream, a stream of its own: the read of chunk k+1 overlaps
// the copy of chunk k, and both overlap whatever the compute stream runs (a
// compaction on another thread).
//
// luda_write_outputs: output tables -> files.  The reference writes each
// output on a pthread of its own with fwrite and fsync (db/db_impl.cc:55-65,
// 1490-1565).  Here the result's bytes come back from HBM in one copy into a
// pinned buffer and up to 16 threads write and sync the files.
//
// luda_encode_edit / luda_encode_compaction_edit / luda_log_record: the
// VersionEdit serialisation (db/version_edit.cc:41-89) and the log record
// framing of the MANIFEST (db/log_writer.cc:35-105, db/log_format.h).
#include <errno.h>
#include <fcntl.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <new>
#include <set>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "engine.h"
#include "luda.h"

struct IoState {
  std::mutex mu;
  int device = 0;
  hipStream_t s = nullptr;
  u8* stage[2] = {nullptr, nullptr};  // pinned read chunks
  hipEvent_t done[2] = {nullptr, nullptr};
  bool pending[2] = {false, false};
  u8* wbuf = nullptr;                 // pinned output staging
  u64 wcap = 0;
};

void io_close(IoState* io) {
  if (!io) return;
  if (io->s) hipStreamSynchronize(io->s);
  for (int i = 0; i < 2; i++) {
    if (io->stage[i]) hipHostFree(io->stage[i]);
    if (io->done[i]) hipEventDestroy(io->done[i]);
  }
  if (io->wbuf) hipHostFree(io->wbuf);
  if (io->s) hipStreamDestroy(io->s);
  delete io;
}

namespace {

constexpr u64 kChunk = 8ull << 20;

struct IoFail {
  int code;
  std::string msg;
};

#define IOCHK(x)                                                                                 \
  do {                                                                                           \
    hipError_t e__ = (x);                                                                        \
    if (e__ != hipSuccess) throw IoFail{LUDA_IO_ERROR, std::string(#x) + ": " + hipGetErrorString(e__)}; \
  } while (0)

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

u64 align_up(u64 x, u64 a) { return (x + a - 1) / a * a; }

template <typename F>
int io_guarded(char* err, size_t errlen, F fn) {
  try {
    fn();
    return LUDA_OK;
  } catch (const IoFail& f) {
    if (err && errlen) snprintf(err, errlen, "%s", f.msg.c_str());
    return f.code;
  } catch (const std::bad_alloc&) {
    if (err && errlen) snprintf(err, errlen, "host memory exhausted");
  } catch (const std::exception& x) {
    if (err && errlen) snprintf(err, errlen, "internal error: %s", x.what());
  } catch (...) {
    if (err && errlen) snprintf(err, errlen, "internal error");
  }
  return LUDA_IO_ERROR;
}

// The context's I/O state, created on first use (caller holds nothing).
IoState* io_of(luda_ctx* c) {
  static std::mutex create_mu;
  std::lock_guard<std::mutex> lk(create_mu);
  IoState** slot = ctx_io_slot(c);
  if (*slot) return *slot;
  IoState* io = new IoState;
  io->device = ctx_device_of(c);
  hipError_t e = hipSetDevice(io->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&io->s, hipStreamNonBlocking);
  for (int i = 0; i < 2 && e == hipSuccess; i++) {
    e = hipHostMalloc(reinterpret_cast<void**>(&io->stage[i]), kChunk, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&io->done[i], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    io_close(io);
    throw IoFail{LUDA_IO_ERROR, std::string("I/O stream set-up: ") + hipGetErrorString(e)};
  }
  *slot = io;
  return io;
}

// pread until n bytes or end of file.
void read_full(int fd, u8* dst, u64 n, u64 off, const char* path) {
  u64 got = 0;
  while (got < n) {
    const ssize_t r = pread(fd, dst + got, n - got, (off_t)(off + got));
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) throw IoFail{LUDA_IO_ERROR, std::string(path) + ": short read"};
    got += (u64)r;
  }
}

void write_full(int fd, const u8* src, u64 n, const char* path) {
  u64 put = 0;
  while (put < n) {
    const ssize_t r = write(fd, src + put, n - put);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) throw IoFail{LUDA_IO_ERROR, std::string(path) + ": write failed"};
    put += (u64)r;
  }
}

void read_files(luda_ctx* c, const char* const* paths, const u64* numbers, u32 n, luda_file_set** out) {
  const double t0 = now_ms();
  IoState* io = io_of(c);
  std::lock_guard<std::mutex> lk(io->mu);
  IOCHK(hipSetDevice(io->device));
  std::vector<int> fds(n, -1);
  std::vector<u64> sizes(n), base(n);
  struct Closer {
    std::vector<int>& f;
    ~Closer() {
      for (int fd : f)
        if (fd >= 0) close(fd);
    }
  } closer{fds};
  u64 total = 0, span = 0;
  for (u32 i = 0; i < n; i++) {
    fds[i] = open(paths[i], O_RDONLY);
    struct stat sb;
    if (fds[i] < 0 || fstat(fds[i], &sb) != 0) throw IoFail{LUDA_IO_ERROR, std::string(paths[i]) + ": cannot open"};
    sizes[i] = (u64)sb.st_size;
    base[i] = span;
    total += sizes[i];
    // each table followed by 16 spare bytes and 256-aligned, as the engine's
    // own upload buffer lays them out
    span += align_up(sizes[i] + 16, 256);
  }
  luda_file_set* set = (luda_file_set*)calloc(1, sizeof(luda_file_set));
  if (!set) throw std::bad_alloc();
  set->files = (luda_file*)calloc(n ? n : 1, sizeof(luda_file));
  if (!set->files) {
    free(set);
    throw std::bad_alloc();
  }
  void* dbuf = nullptr;
  hipError_t e = hipMalloc(&dbuf, std::max<u64>(span, 256));
  if (e != hipSuccess) {
    free(set->files);
    free(set);
    throw IoFail{LUDA_IO_ERROR, std::string("device buffer for the input files: ") + hipGetErrorString(e)};
  }
  try {
    u32 k = 0;
    for (u32 i = 0; i < n; i++) {
      for (u64 off = 0; off < sizes[i]; off += kChunk) {
        const u32 b = k++ & 1;
        if (io->pending[b]) IOCHK(hipEventSynchronize(io->done[b]));  // its last copy has read it
        const u64 len = std::min<u64>(kChunk, sizes[i] - off);
        read_full(fds[i], io->stage[b], len, off, paths[i]);
        IOCHK(hipMemcpyAsync(static_cast<u8*>(dbuf) + base[i] + off, io->stage[b], len, hipMemcpyHostToDevice, io->s));
        IOCHK(hipEventRecord(io->done[b], io->s));
        io->pending[b] = true;
      }
      set->files[i].data = static_cast<const u8*>(dbuf) + base[i];
      set->files[i].size = sizes[i];
      set->files[i].number = numbers ? numbers[i] : i + 1;
    }
    IOCHK(hipStreamSynchronize(io->s));
    io->pending[0] = io->pending[1] = false;
  } catch (...) {
    hipStreamSynchronize(io->s);
    io->pending[0] = io->pending[1] = false;
    hipFree(dbuf);
    free(set->files);
    free(set);
    throw;
  }
  set->n = n;
  set->device_data = dbuf;
  set->total_bytes = total;
  set->read_ms = now_ms() - t0;
  *out = set;
}

void write_outputs(luda_ctx* c, const luda_result* r, const char* const* paths, int sync) {
  const u8* src = r->data;
  IoState* io = io_of(c);
  std::unique_lock<std::mutex> lk(io->mu);
  if (!src && r->total_bytes) {
    IOCHK(hipSetDevice(io->device));
    if (io->wcap < r->total_bytes) {
      if (io->wbuf) hipHostFree(io->wbuf);
      io->wbuf = nullptr;
      io->wcap = 0;
      IOCHK(hipHostMalloc(reinterpret_cast<void**>(&io->wbuf), r->total_bytes, hipHostMallocDefault));
      io->wcap = r->total_bytes;
    }
    IOCHK(hipMemcpyAsync(io->wbuf, r->device_data, r->total_bytes, hipMemcpyDeviceToHost, io->s));
    IOCHK(hipStreamSynchronize(io->s));
    src = io->wbuf;
  }
  const u32 n = r->num_outputs;
  std::atomic<u32> next{0};
  std::mutex emu;
  std::string first_err;
  auto worker = [&] {
    for (;;) {
      const u32 i = next.fetch_add(1);
      if (i >= n) return;
      try {
        const int fd = open(paths[i], O_WRONLY | O_CREAT | O_TRUNC, 0644);
        if (fd < 0) throw IoFail{LUDA_IO_ERROR, std::string(paths[i]) + ": cannot create"};
        try {
          write_full(fd, src + r->outputs[i].offset, r->outputs[i].file_size, paths[i]);
          if (sync && fsync(fd) != 0) throw IoFail{LUDA_IO_ERROR, std::string(paths[i]) + ": fsync failed"};
        } catch (...) {
          close(fd);
          throw;
        }
        if (close(fd) != 0) throw IoFail{LUDA_IO_ERROR, std::string(paths[i]) + ": close failed"};
      } catch (const IoFail& f) {
        std::lock_guard<std::mutex> g(emu);
        if (first_err.empty()) first_err = f.msg;
      }
    }
  };
  const u32 nt = std::max<u32>(1, std::min<u32>(16, n));
  std::vector<std::thread> th;
  for (u32 t = 1; t < nt; t++) th.emplace_back(worker);
  worker();
  for (auto& t : th) t.join();
  if (!first_err.empty()) throw IoFail{LUDA_IO_ERROR, first_err};
}

// ---------------------------------------------------------------------------
// VersionEdit and log records

struct Sink {
  uint8_t* dst;
  u64 cap;
  u64 len = 0;
  void put(const void* p, u64 n) {
    if (len + n <= cap && n) memcpy(dst + len, p, n);
    len += n;
  }
  void byte(uint8_t b) { put(&b, 1); }
  // PutVarint32/64 (util/coding.cc:48-75)
  void varint(u64 v) {
    uint8_t b[10];
    int k = 0;
    while (v >= 128) {
      b[k++] = (uint8_t)(v | 128);
      v >>= 7;
    }
    b[k++] = (uint8_t)v;
    put(b, k);
  }
  // PutLengthPrefixedSlice (util/coding.cc:82-85)
  void slice(const luda_slice& s) {
    varint(s.size);
    put(s.data, s.size);
  }
};

// VersionEdit tags (db/version_edit.cc:14-24)
enum : u32 {
  kComparator = 1,
  kLogNumber = 2,
  kNextFileNumber = 3,
  kLastSequence = 4,
  kCompactPointer = 5,
  kDeletedFile = 6,
  kNewFile = 7,
  kPrevLogNumber = 9,
};

// VersionEdit::EncodeTo (db/version_edit.cc:41-89), field by field in its
// order; `deleted` is the edit's deleted-file set.
void encode_edit(Sink& o, const luda_edit& e, const std::set<std::pair<int, u64>>& deleted) {
  if (e.has & LUDA_EDIT_HAS_COMPARATOR) {
    o.varint(kComparator);
    o.slice(e.comparator);
  }
  if (e.has & LUDA_EDIT_HAS_LOG_NUMBER) {
    o.varint(kLogNumber);
    o.varint(e.log_number);
  }
  if (e.has & LUDA_EDIT_HAS_PREV_LOG_NUMBER) {
    o.varint(kPrevLogNumber);
    o.varint(e.prev_log_number);
  }
  if (e.has & LUDA_EDIT_HAS_NEXT_FILE) {
    o.varint(kNextFileNumber);
    o.varint(e.next_file_number);
  }
  if (e.has & LUDA_EDIT_HAS_LAST_SEQUENCE) {
    o.varint(kLastSequence);
    o.varint(e.last_sequence);
  }
  for (u32 i = 0; i < e.num_compact_pointers; i++) {
    o.varint(kCompactPointer);
    o.varint((u32)e.compact_pointers[i].level);
    o.slice(e.compact_pointers[i].key);
  }
  for (const auto& d : deleted) {
    o.varint(kDeletedFile);
    o.varint((u32)d.first);
    o.varint(d.second);
  }
  for (u32 i = 0; i < e.num_new_files; i++) {
    const luda_edit_file& f = e.new_files[i];
    o.varint(kNewFile);
    o.varint((u32)f.level);
    o.varint(f.number);
    o.varint(f.file_size);
    o.slice(f.smallest);
    o.slice(f.largest);
  }
}

std::set<std::pair<int, u64>> deleted_set(const luda_edit& e) {
  std::set<std::pair<int, u64>> d;
  for (u32 i = 0; i < e.num_deleted; i++) d.insert({e.deleted[i].level, e.deleted[i].number});
  return d;
}

bool edit_ok(const luda_edit& e) {
  if (e.num_compact_pointers && !e.compact_pointers) return false;
  if (e.num_deleted && !e.deleted) return false;
  if (e.num_new_files && !e.new_files) return false;
  return true;
}

// CRC-32C (Castagnoli, reflected 0x82F63B78) of the log header: crc32c::Value
// and Extend (util/crc32c.cc:276-377) with a byte-wise table; mask as in
// util/crc32c.h:29-32.
struct HostCrc {
  u32 t[256];
  HostCrc() {
    for (u32 i = 0; i < 256; i++) {
      u32 c = i;
      for (int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ 0x82F63B78u : (c >> 1);
      t[i] = c;
    }
  }
  u32 extend(u32 crc, const uint8_t* p, u64 n) const {
    u32 c = ~crc;
    for (u64 i = 0; i < n; i++) c = t[(c ^ p[i]) & 0xff] ^ (c >> 8);
    return ~c;
  }
};

u32 mask_crc(u32 crc) { return ((crc >> 15) | (crc << 17)) + 0xa282ead8u; }

}  // namespace

extern "C" {

int luda_read_files(luda_ctx* c, const char* const* paths, const uint64_t* numbers, uint32_t n, luda_file_set** out,
                    char* err, size_t errlen) {
  if (!c || !out || (n && !paths)) return LUDA_INVALID_ARGUMENT;
  *out = nullptr;
  return io_guarded(err, errlen, [&] { read_files(c, paths, numbers, n, out); });
}

void luda_free_files(luda_ctx* c, luda_file_set* set) {
  if (!set) return;
  if (set->device_data) {
    if (c) hipSetDevice(ctx_device_of(c));
    hipFree(set->device_data);
  }
  free(set->files);
  free(set);
}

int luda_write_outputs(luda_ctx* c, const luda_result* res, const char* const* paths, int sync, char* err,
                       size_t errlen) {
  if (!c || !res || (res->num_outputs && !paths)) return LUDA_INVALID_ARGUMENT;
  if (!res->data && res->total_bytes && !res->device_data) return LUDA_INVALID_ARGUMENT;
  return io_guarded(err, errlen, [&] { write_outputs(c, res, paths, sync); });
}

int luda_encode_edit(const luda_edit* e, uint8_t* dst, uint64_t cap, uint64_t* len) {
  if (!e || !len || (cap && !dst) || !edit_ok(*e)) return LUDA_INVALID_ARGUMENT;
  return io_guarded(nullptr, 0, [&] {
    Sink o{dst, cap};
    encode_edit(o, *e, deleted_set(*e));
    *len = o.len;
    if (o.len > cap) throw IoFail{LUDA_INVALID_ARGUMENT, "buffer too small"};
  });
}

int luda_encode_compaction_edit(const luda_job* job, const luda_result* res, const uint64_t* output_numbers,
                                const luda_edit* base, uint8_t* dst, uint64_t cap, uint64_t* len) {
  if (!job || !res || !len || (res->num_outputs && !output_numbers) || (cap && !dst)) return LUDA_INVALID_ARGUMENT;
  if (base && !edit_ok(*base)) return LUDA_INVALID_ARGUMENT;
  return io_guarded(nullptr, 0, [&] {
    luda_edit e;
    memset(&e, 0, sizeof(e));
    if (base) e = *base;
    // Compaction::AddInputDeletions (db/version_set.cc:1527-1533): inputs_[w]
    // come from level + w.
    std::set<std::pair<int, u64>> del = base ? deleted_set(*base) : std::set<std::pair<int, u64>>();
    for (int w = 0; w < 2; w++)
      for (u32 i = 0; i < job->num_inputs[w]; i++) del.insert({job->level + w, job->inputs[w][i].number});
    // InstallCompactionResults (db/db_impl.cc:1048-1053): outputs at level + 1
    std::vector<luda_edit_file> nf(e.new_files, e.new_files + e.num_new_files);
    for (u32 i = 0; i < res->num_outputs; i++) {
      const luda_output& o = res->outputs[i];
      luda_edit_file f;
      memset(&f, 0, sizeof(f));
      f.level = job->level + 1;
      f.number = output_numbers[i];
      f.file_size = o.file_size;
      f.smallest = luda_slice{res->keys + o.smallest_offset, o.smallest_size};
      f.largest = luda_slice{res->keys + o.largest_offset, o.largest_size};
      nf.push_back(f);
    }
    e.new_files = nf.data();
    e.num_new_files = (u32)nf.size();
    Sink o{dst, cap};
    encode_edit(o, e, del);
    *len = o.len;
    if (o.len > cap) throw IoFail{LUDA_INVALID_ARGUMENT, "buffer too small"};
  });
}

int luda_log_record(uint64_t* block_offset, const uint8_t* payload, uint64_t n, uint8_t* dst, uint64_t cap,
                    uint64_t* len) {
  if (!block_offset || !len || (n && !payload) || (cap && !dst)) return LUDA_INVALID_ARGUMENT;
  constexpr u64 kBlockSize = 32768, kHeaderSize = 7;  // db/log_format.h
  static const HostCrc crc;
  Sink o{dst, cap};
  u64 boff = *block_offset % kBlockSize;
  u64 left = n;
  const uint8_t* p = payload;
  bool begin = true;
  // log::Writer::AddRecord (db/log_writer.cc:35-80): an empty record still
  // emits one zero-length fragment
  do {
    const u64 leftover = kBlockSize - boff;
    if (leftover < kHeaderSize) {
      static const uint8_t zeros[kHeaderSize] = {0};
      o.put(zeros, leftover);  // block trailer
      boff = 0;
    }
    const u64 avail = kBlockSize - boff - kHeaderSize;
    const u64 frag = left < avail ? left : avail;
    const bool end = left == frag;
    const uint8_t type = begin && end ? 1 : (begin ? 2 : (end ? 4 : 3));  // FULL, FIRST, LAST, MIDDLE
    // EmitPhysicalRecord (db/log_writer.cc:82-105): masked crc of type and
    // data, 2-byte length, type
    const u32 cv = mask_crc(crc.extend(crc.extend(0, &type, 1), p, frag));
    uint8_t h[kHeaderSize] = {(uint8_t)cv,         (uint8_t)(cv >> 8), (uint8_t)(cv >> 16), (uint8_t)(cv >> 24),
                              (uint8_t)(frag & 0xff), (uint8_t)(frag >> 8), type};
    o.put(h, kHeaderSize);
    o.put(p, frag);
    boff += kHeaderSize + frag;
    p += frag;
    left -= frag;
    begin = false;
  } while (left > 0);
  *len = o.len;
  if (o.len > cap) return LUDA_INVALID_ARGUMENT;
  *block_offset = boff % kBlockSize;  // the file length mod the block size
  return LUDA_OK;
}

}  // extern "C"

// engine.h — structures shared by the host driver (engine.hip) and the kernel
// units (k_*.hip), and the host-side launchers each unit exports.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "cuts.h"
#include "dev_common.h"

// ---------------------------------------------------------------------------
// per-kernel timing with HIP events on the engine's stream

struct KTimer {
  hipStream_t s = nullptr;
  bool on = false;
  struct Rec {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  hipEvent_t get() {
    if (used == pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool.push_back(e);
    }
    return pool[used++];
  }
  const char* cur = nullptr;
  hipEvent_t cur_a = nullptr;
  std::vector<std::string> only;  // when non-empty, time just these kernels
  void begin(const char* name) {
    cur_a = nullptr;
    if (!on) return;
    if (!only.empty()) {
      bool hit = false;
      for (auto& o : only) hit |= (o == name);
      if (!hit) return;
    }
    cur = name;
    cur_a = get();
    if (cur_a) hipEventRecord(cur_a, s);
  }
  void end() {
    if (!on || !cur_a) return;
    hipEvent_t b = get();
    if (!b) return;
    hipEventRecord(b, s);
    recs.push_back({cur, cur_a, b});
  }
  void reset() {
    recs.clear();
    used = 0;
  }
  ~KTimer() {
    for (auto e : pool) hipEventDestroy(e);
  }
};

// Throws the engine's error for a launch the runtime refused (bad shape or
// LDS size): nothing after it is enqueued, so no later kernel reads buffers
// the refused one should have written.
[[noreturn]] void launch_refused(const char* name, hipError_t e);

#define KLAUNCH(T, name, kern, grid, block, lds, ...)                                     \
  do {                                                                                    \
    (T).begin(name);                                                                      \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), (lds), (T).s, __VA_ARGS__);       \
    const hipError_t klaunch_e = hipGetLastError();                                       \
    if (klaunch_e != hipSuccess) launch_refused(name, klaunch_e);                         \
    (T).end();                                                                            \
  } while (0)

// ---------------------------------------------------------------------------
// what io.hip needs of a context: its device and the slot for its I/O state
// (created on first use, closed with the context)

struct luda_ctx;
struct IoState;
int ctx_device_of(luda_ctx* c);
IoState** ctx_io_slot(luda_ctx* c);
void io_close(IoState* io);

// ---------------------------------------------------------------------------
// input side

struct FileDesc {
  const u8* base;  // device address of the table's first byte
  u64 size;
};
struct FileInfo {
  u64 ioff;   // index block handle
  u64 isize;
  u32 nr;     // index restarts
  u32 pad;
};
struct BlockDesc {
  u64 addr;   // device address of the data block's first byte
  u32 size;   // BlockHandle::size (without the 5-byte trailer)
  u32 file;
};

// One input table whose decoded entries are resident: entries [first,
// first + count) of the arrays below go to [dst, dst + count) of the job.
struct DecodedSrc {
  const u64* P;    // plane words, W planes of `stride`
  u64 stride;
  u32 W;
  u32 pad;
  const u32* klen;
  const u64* tag;
  const u64* vptr;  // device addresses of the values in the table itself
  const u32* vlen;
  u64 first;
  u64 count;
  u64 dst;
};

// Two adjacent runs of an index array merged into the same positions:
// [a0, a1) and [a1, b1); tile0 = first thread-tile of this pair.
struct MergePair {
  u32 a0, a1, b1, tile0;
};

// Parameters of the drop rule (db/db_impl.cc:1113-1147), IsBaseLevelForKey
// (db/version_set.cc:1535-1554) and ShouldStopBefore (:1556-1579).
struct DropParams {
  u64 S;            // smallest_snapshot
  u64 limit;        // MaxGrandParentOverlapBytes = 10 * max_file_size
  u32 ngp;          // grandparents
  u32 ndeep;
  const u64* gp_words;   // ngp x W: largest user key, first 8W bytes
  const u32* gp_len;     // user-key length of each largest key
  const u64* gp_tag;
  const u64* gp_prefix;  // ngp + 1 running sums of file sizes
  const u64* deep_words; // ndeep x 2 x W: smallest, largest user keys
  const u32* deep_len;   // ndeep x 2
  u32 lvl_begin[8], lvl_end[8];  // deeper files of level L at [lvl_begin[L], lvl_end[L])
  u32 first_level;       // level + 2
  u32 pad;
};

// The stream the table encoder consumes: kept entries in output order.
struct Kept {
  Planes K;
  u64* vptr;    // device address of the value
  u32* vlen;
  u32* forced;  // 1 where a new table must start before this entry
};

struct EncParams {
  u64 block_size;
  u64 max_file_size;
  u32 interval;     // block_restart_interval
  u32 bloom_bits;   // 0 = no filter block
  u32 bloom_k;
  u32 kw;           // bytes of one stored internal key (8W + 8)
};

// Per output table, read back by the host.
struct FileMeta {
  u64 offset;       // in the output buffer
  u64 size;
  u64 nent;
  u32 slen, llen;   // smallest / largest internal key lengths
};

// Per output block and per output table working arrays (see k_encode.hip).
struct EncBufs {
  // per kept entry
  u32* shared;   // shared-prefix length with the previous kept key
  u32* szr;      // encoded size as a restart entry
  u32* sznr;     // encoded size as a prefix-compressed entry
  u32* bnext;    // first entry after the block that would start here
  u32* bbytes;   // raw size of that block
  u32* bs;       // 1 where a block starts
  // per output block (nbo + 1 entries where a sentinel is needed)
  u32* ob_first;
  u32* ob_n;
  u32* ob_bytes;
  u32* ob_forced;  // 1 where ShouldStopBefore starts a table at this block
  // the table that would start at each block (k_table_next)
  u16* tnext;      // its end, as a distance (0: past the staged span)
  u64* torel;      // the block's offset in its tile of kTableTile blocks
  u64* tdend;      // its data bytes
  u32* tnw;        // its filter windows
  u64* tbytes;     // per tile: bytes of its blocks
  u32* tforced;    // per tile: its first forced block, or ~0
  u32* ob_file;
  u64* ob_off;     // offset inside its table
  u64* ob_isz;     // offset of its index entry in the index block
  // per output table (nf + 1 entries)
  u32* f_first;    // first block; f_first[nf] = nbo
  u32* f_win;      // first filter window (windows are numbered across tables)
  u64* f_dend;     // data end offset
  u64* f_fbytes;   // filter data bytes
  u64* f_ibytes;   // index block size
  u64* f_fbase;    // where the filter data starts in the table (its data end)
  u64* f_off;      // table size, then (scanned in place) table offset in the output
  FileMeta* meta;
  u8* keys;        // 2 x kw bytes per table: smallest, largest
  // per filter window (nw + 1 entries)
  u32* w_keys;
  u32* w_first;
  u32* w_file;     // its table (windows with keys only)
  u64* w_fsz;      // filter bytes, then (scanned in place) offset in its table's filter data
};

// ---------------------------------------------------------------------------
// launchers

// A context's single-pass scans: one look-back word per tile of the largest
// scan (zeroed when allocated, never shared with other data) and the epoch of
// the last launch, which tags the words it writes.
struct ScanState {
  u64* words = nullptr;
  u64 tiles = 0;
  u32 epoch = 0;  // 1 .. 2^14 - 1; the words are zeroed again when it wraps
  u32* ticket = nullptr;  // the word after the tiles' words: next tile number of the running scan, 0 between launches
  bool tickets = false;   // tiles by ticket (worker slots), else by blockIdx
};
struct ScanArgs {
  u64* words;
  u32* ticket;  // null: tile t is workgroup t
  u32 epoch;
  u32 cap;    // with need: a total above cap is stored to *need and replaced by 0
  u32* need;
  DevStatus* st;
};
struct ScanGuard {
  u32 cap;
  u32* need;
};
u64 scan_tiles_for(u64 n_cap);
// The words for scans of up to n_cap items (the stream is idle when it grows).
hipError_t scan_reserve(ScanState& S, u64 n_cap, hipStream_t s);
// The words for a look-back over `tiles` tiles of any size.
hipError_t scan_reserve_tiles(ScanState& S, u64 tiles, hipStream_t s);
void scan_release(ScanState& S);
// Look-back words the caller owns (`tiles` u64 at `words`), zeroed on the
// stream; such a state is never grown or released by the scans.
hipError_t scan_attach(ScanState& S, u64* words, u64 words_n, hipStream_t s);  // words_n - 1 tiles, then the ticket
u32 scan_next_epoch(ScanState& S, hipStream_t s);
[[noreturn]] void throw_scan_capacity(u64 need, u64 have);
// Exclusive scan of n = (n_dev ? *n_dev : 0) + n_add items (n <= n_cap);
// total to *total (may be null).  In place is allowed.  With a guard, a
// total above g->cap is reported (ERR_OVERFLOW, *g->need) and stored as 0.
void scan_u32(KTimer& T, ScanState& S, const u32* in, u32* out, const u32* n_dev, u32 n_add, u32 n_cap, u32* total,
              DevStatus* st, const ScanGuard* g = nullptr);
void scan_u64(KTimer& T, ScanState& S, const u64* in, u64* out, const u32* n_dev, u32 n_add, u32 n_cap, u64* total,
              DevStatus* st);
void scan_u32_to_u64(KTimer& T, ScanState& S, const u32* in, u64* out, const u32* n_dev, u32 n_add, u32 n_cap,
                     u64* total, DevStatus* st);
void fill_u32(KTimer& T, u32* p, u32 v, u64 n);
void fill_u8(KTimer& T, u8* p, u8 v, u64 n);
void iota_u32(KTimer& T, u32* p, u64 n);

// k_decode.hip
// k_stage.hip: src/dst on the host side are device addresses of pinned host
// memory (hipHostGetDevicePointer).  stage_in copies `bytes` from src to dst
// and, with init, the status; stage_out copies each segment (device src) to
// dst + off.
struct StageSeg {
  const void* src;
  u64 off;
  u64 bytes;
};
// n words at p set to v by the same launch (device memory the staging itself
// does not touch)
struct StageFill {
  u32* p;
  u64 n;
  u32 v;
};
void launch_stage_in(KTimer& T, const u8* src, u8* dst, u64 bytes, const DevStatus* init, DevStatus* st,
                     StageFill fill = StageFill{nullptr, 0, 0});
// k_cuts.hip: the output block cuts by tiles of kCutTile entries (staged with
// kCutSpan more), for chains with at most kCutCand entry points per tile.
// ftab: kCutCand words per tile (exit | blocks << 12); cand, entry, base:
// one word per tile.
constexpr u32 kCutTile = cuts::kTile;
constexpr u32 kCutSpan = cuts::kSpan;
constexpr u32 kCutCand = cuts::kCand;
struct CutBufs {
  u32* ftab;
  u32* cand;
  u32* entry;
  u32* base;
};
inline u32 cut_tiles_for(u64 m_cap) { return (u32)(m_cap / kCutTile + 1); }
void run_block_cuts(KTimer& T, Kept k, const u32* m_dev, u32 m_cap, EncParams ep, EncBufs b, CutBufs cb, u32 nbo_cap,
                    DevStatus* st);

// Completion flag of a stage_out: flag (device address of host memory) gets
// seq once every segment is in place; ctr is a device word, zero between
// launches.  flag null: no flag.
struct StageDone {
  u64* flag;
  u32* ctr;
  u64 seq;
};
void launch_stage_out(KTimer& T, u8* dst, StageSeg s0, StageSeg s1, StageSeg s2, StageDone done);

// The status (from init) and the descriptors (from src) of a read of tables,
// both in pinned host memory; footers; blk_base = each table's first data
// block, the total at blk_base[nfiles].
void launch_files_in(KTimer& T, const FileDesc* src, const DevStatus* init, FileDesc* files, u32 nfiles,
                     FileInfo* info, u32* blk_base, DevStatus* st);
// k_verify.hip: every block's masked CRC (data blocks from launch_index's
// descriptors, then per table the index, metaindex and named meta blocks);
// a failure sets ERR_CORRUPT / SITE_CHECKSUM and the lowest bad table in
// st->bad_file.
void launch_verify(KTimer& T, const FileDesc* files, u32 nfiles, const BlockDesc* blks, u32 nb_cap, DevStatus* st);
void launch_index(KTimer& T, const FileDesc* files, const FileInfo* info, const u32* blk_base, u32 nfiles,
                  u32 nb_cap, BlockDesc* blks, u32* blk_nent, DevStatus* st);
// blk_ebase = exclusive scan of blk_nent over st->n_blocks_in blocks,
// st->n_entries, and file_ebase[f] = the first entry of table f (f <= nfiles)
// launch_index that also scans the blocks' entry counts: each block's first
// entry (blk_ebase), each table's (file_ebase, nfiles + 1 of them) and the
// total (st->n_entries), by a look-back over tiles of 256 blocks; the look-
// back words must hold index_scan_tiles(nb_cap)
u64 index_scan_tiles(u32 nb_cap);
void launch_index_entries(KTimer& T, ScanState& S, const FileDesc* files, const FileInfo* info, const u32* blk_base,
                          u32 nfiles, u32 nb_cap, BlockDesc* blks, u32* blk_nent, u32* blk_ebase, u32* file_ebase,
                          DevStatus* st);
// A decoder launched behind phase A before the host has read its counts:
// it decodes only when they fit the launch (no error, at most n_cap entries,
// restart keys of at most max_rulen bytes, at most max_nr restart points a
// block: the max_nr of the launch), and then over st->n_blocks_in blocks;
// otherwise every workgroup returns before its first write.
struct DecodeGuard {
  u32 on;
  u32 n_cap;
  u32 max_rulen;
};
// groups: wavefronts of the decoder, 0 = the default
void launch_decode(KTimer& T, const FileDesc* files, const BlockDesc* blks, u32 nb, const u32* blk_ebase, Planes K,
                   u64* vptr, u32* vlen, DevStatus* st, u32 max_blk, u32 max_nr, u32 max_ent, u32 groups = 0,
                   DecodeGuard guard = DecodeGuard{0, 0, 0});
// LDS of k_decode_entries for these maxima (over 64 KiB: it cannot run)
u64 decode_lds_bytes(u32 max_blk, u32 max_nr);
void launch_ingest(KTimer& T, u64 n, const u8* keys, const u64* key_off, const u8* vals, const u64* val_off, Kept out,
                   DevStatus* st);
// entries 0 .. n-1 from resident decoded sources (src sorted by dst)
void launch_gather_decoded(KTimer& T, const DecodedSrc* src, u32 nsrc, u64 n, Planes K, u64* vptr, u32* vlen);

// k_merge.hip
// One merge round; in == nullptr reads the children in place (first round,
// key planes of at most 8 words), and check verifies each child is sorted.
// splits: one u32 per merge tile of the round (working memory).
void launch_merge(KTimer& T, const MergePair* pairs, u32 npairs, u32 total_tiles, const u32* in, u32* out, Planes K,
                  bool check, u32* splits, DevStatus* st);
// The last round (one pair; W <= 8) fused with the drop rule, the kept-count
// scan and the gather into the kept stream; with grandparents also gcnt,
// kpos and the crossing slots (cleared to all ones before) for launch_stops.
// Sets st->n_kept.
void launch_merge_select(KTimer& T, ScanState& S, const MergePair* d_pair, MergePair pr, const u32* in, bool check,
                         Planes K, const u64* vptr, const u32* vlen, const DropParams& dp, Kept out, u32* gcnt,
                         u32* kpos, u32* slot, u32* splits, DevStatus* st);
u32 merge_items_per_tile(u32 W);  // outputs per merge tile for key planes of W words
// also writes iota[e] = e, the merge's starting permutation
void launch_check_sorted(KTimer& T, Planes K, u64 n, const u32* child_start, u32 nchild, u32* iota, DevStatus* st);

// k_select.hip
// events[0..ngp]: the grandparent-crossing slots, cleared here and filled by
// the keep scan (events[g] = the entry where the grandparent count reaches g)
void launch_drop(KTimer& T, const u32* perm, u32 n, Planes K, const DropParams& dp, u32* keep, u32* gcnt, u32* events,
                 DevStatus* st);
// kpos = exclusive scan of keep, n_kept; with grandparents also the crossing
// slots (events[1..ngp]) that launch_stops walks
void scan_keep(KTimer& T, ScanState& S, const u32* keep, u32* kpos, u32 n, const DropParams& dp, const u32* gcnt,
               u32* events, u32 ev_cap, DevStatus* st);
void launch_gather(KTimer& T, const u32* perm, u32 n, const u32* keep, const u32* kpos, Planes K, const u64* vptr,
                   const u32* vlen, Kept out);
void launch_stops(KTimer& T, const u32* gcnt, const u32* kpos, const u32* keep, u32 n, const DropParams& dp,
                  u32* events, u32 ev_cap, Kept out, DevStatus* st);

// k_orbit.hip: marks in flag[] the nodes reached from node 0 by nxt[].
struct OrbitPlan {
  u32 C;       // chunk size
  u32 Wc;      // window per chunk (max step bound)
  u32 G;       // composition group size
  u32 levels;  // composition levels above the chunk level
  u64 n_cap;
  std::vector<u64> level_items;   // capacity of items per level
  std::vector<u64> level_offset;  // offsets (in u16 items of width Wc) of each level's tables
  u64 table_entries;              // total u16 entries
  u64 entry_entries;              // total u32 entries
};
OrbitPlan plan_orbit(u64 n_cap, u32 C, u32 Wc, u32 G);
void run_orbit(KTimer& T, const OrbitPlan& p, const u32* nxt, const u32* n_dev, const u32* step_dev, u32* flag,
               u16* ftab, u32* entries, DevStatus* st);

// k_encode.hip
void launch_entry_sizes(KTimer& T, Kept k, const u32* m_dev, u32 m_cap, EncBufs b, DevStatus* st);
void launch_bnext(KTimer& T, Kept k, const u32* m_dev, u32 m_cap, EncParams ep, EncBufs b, DevStatus* st);
// scans with the per-block and per-table steps fused in (n_blocks_out with
// the block table, block offsets O, the forced-block list, n_files_out with
// f_first, the window offsets f_win, the table offsets f_off)
void scan_blocks(KTimer& T, ScanState& S, Kept k, const u32* m_dev, u32 m_cap, EncBufs b, u32 nbo_cap,
                 DevStatus* st);
// Table cuts (f_first, n_files_out), data_bytes, each table's data end and
// first filter window (f_dend, f_win, n_windows); f_off[n_files_out] = 0 for
// the table offset scan.  Two launches: the table that would start at each
// block, then the chain of those from block 0 in one workgroup.
constexpr u32 kTableTile = 1024;
void launch_cut_tables(KTimer& T, const u32* nbo_dev, u32 nbo_cap, EncParams ep, EncBufs b, u32 nf_cap, u32 nw_cap,
                       bool size_cuts, DevStatus* st);
// Per table: ob_file, ob_off, ob_isz, the windows (w_*), f_fbytes, f_fbase,
// f_ibytes and the table's size in f_off.
// each table's blocks, index entries and filter windows, and (look-back over
// the tables) its offset in the output; the output's size into st->out_bytes
void launch_table_layout(KTimer& T, ScanState& S, Kept k, const u32* nf_dev, u32 nf_cap, EncParams ep, EncBufs b,
                         DevStatus* st);
// groups: workgroups of the block writer, 0 = one per CU
void launch_write_blocks(KTimer& T, Kept k, const u32* nbo_dev, u32 nbo_cap, EncParams ep, EncBufs b, u8* out,
                         u64 out_cap, DevStatus* st, u32 groups = 0);
// vaddr[q] = device address of kept entry q's value in the written tables
void launch_value_addrs(KTimer& T, Kept k, const u32* nbo_dev, u32 nbo_cap, EncParams ep, EncBufs b, u8* out,
                        u64* vaddr);
void launch_write_filters(KTimer& T, Kept k, const u32* nw_dev, u32 nw_cap, const u32* nf_dev, EncParams ep, EncBufs b,
                          u8* out, u64 out_cap, DevStatus* st);
void launch_write_tail(KTimer& T, Kept k, const u32* nf_dev, u32 nf_cap, EncParams ep, EncBufs b, u8* out,
                       u64 out_cap, DevStatus* st);

// k_read.hip: ksz[e] = internal-key bytes of entry e (e < n), ksz[n] = vlen[n]
// = 0; then the entries to plain arrays
void launch_entry_out_sizes(KTimer& T, Planes K, u64 n, u32* ksz, u32* vlen);
void launch_emit(KTimer& T, Planes K, const u64* vptr, const u32* vlen, u64 n, const u64* key_off, const u64* val_off,
                 u8* keys, u8* vals);

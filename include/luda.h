/*
 * luda.h — C-ABI of the MI355X LSM compaction engine.
 *
 * The engine is a drop-in for the compaction step of the LevelDB 1.22 fork in
 * xpsair/luda.  A caller that holds a leveldb::Compaction hands its input
 * tables to luda_compact() and gets back finished table images plus the
 * metadata that DBImpl::InstallCompactionResults puts into the VersionEdit.
 * The output bytes are identical to what the reference's CPU
 * DBImpl::DoCompactionWork (db/db_impl.cc:1060-1224) writes for the same
 * inputs.  Nothing here uses C++ types; every pointer is a plain buffer.
 *
 * Reference interfaces replaced (all paths relative to the reference tree):
 *   luda_compact        DBImpl::DoCompactionWork, db/db_impl.cc:1060-1224
 *                       (iterator through the last FinishCompactionOutputFile,
 *                       :1079-1197), with Compaction::ShouldStopBefore and
 *                       Compaction::IsBaseLevelForKey, db/version_set.cc:1535-1579
 *   luda_build_tables   BuildTable, db/builder.cc:17-77 (one table per call when
 *                       max_file_size is UINT64_MAX), and TableBuilder,
 *                       table/table_builder.cc:94-254
 *   luda_init/shutdown  gpu::HostAndDeviceMemory, cuda/decode_kv.cu:43-168
 *                       (one engine context per DBImpl, db/db_impl.h:215)
 */
#ifndef LUDA_H_
#define LUDA_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, one per leveldb::Status kind (include/leveldb/status.h:98-105).
 * Any code other than LUDA_OK means: run the unchanged CPU compaction. */
#define LUDA_OK 0
#define LUDA_NOT_FOUND 1
#define LUDA_CORRUPTION 2
#define LUDA_NOT_SUPPORTED 3
#define LUDA_INVALID_ARGUMENT 4
#define LUDA_IO_ERROR 5

/* luda_job.flags */
#define LUDA_JOB_INPUTS_ON_DEVICE 0x1u /* luda_file.data are device pointers */
#define LUDA_JOB_KEEP_ON_DEVICE 0x2u   /* leave output bytes in HBM only (result.data == NULL) */
/* Keep the output tables' decoded entries (key planes, tags, value addresses
 * in the tables) resident in HBM with the result, until luda_free_result.  A
 * later luda_compact whose inputs are all such tables, passed by their device
 * addresses (LUDA_JOB_INPUTS_ON_DEVICE) to the context that wrote them (its
 * worker slots included; other contexts decode them as usual), takes their
 * entries from there and decodes nothing: the chained level compactions of a
 * served DB.  The
 * reference keeps the decoded entries of flushed level-0 tables in host memory
 * for the next L0->L1 compaction the same way (cuda/cuda_common.h:253-255,
 * 339-355; db/db_impl.cc:146-149,1292-1295).  Also a flag of
 * luda_build_tables. */
#define LUDA_JOB_KEEP_DECODED 0x8u

/* The subset of leveldb::Options (include/leveldb/options.h:33-143) that
 * changes table bytes. */
typedef struct luda_options {
  uint64_t block_size;             /* Options::block_size, options.h:100 (4096) */
  uint64_t max_file_size;          /* Options::max_file_size, options.h:115 (2 MiB) */
  uint32_t block_restart_interval; /* Options::block_restart_interval, options.h:105 (16).  The fork
                                      stores each restart as (offset << 8) | count
                                      (table/block_builder.cc:62-68,89-94), so a restart group of 256 or
                                      more entries spills its count into the offset bits.  The writers
                                      reproduce those bytes for any interval, as the fork's release build
                                      does; tables with such a group cannot be read back, by the fork's
                                      Block::Iter or by luda_compact/luda_verify_tables (CORRUPTION).
                                      Intervals up to 255 always round-trip. */
  uint32_t bloom_bits_per_key;     /* NewBloomFilterPolicy(bits), util/bloom.cc:17; 0 = no filter_policy */
  uint32_t compression;            /* Options::compression; only kNoCompression (0) is supported */
  uint32_t reserved;
} luda_options;

/* One input table: FileMetaData{number, file_size} plus the file's bytes
 * (db/version_edit.h:18-27). */
typedef struct luda_file {
  const uint8_t* data;
  uint64_t size;
  uint64_t number;
} luda_file;

typedef struct luda_slice {
  const uint8_t* data;
  uint64_t size;
} luda_slice;

/* Compaction::grandparents_ entry as read by ShouldStopBefore
 * (db/version_set.cc:1556-1579): the file's largest internal key and size. */
typedef struct luda_grandparent {
  luda_slice largest;
  uint64_t file_size;
} luda_grandparent;

/* A file of a level >= level+2, as read by IsBaseLevelForKey
 * (db/version_set.cc:1535-1554).  Files may come in any level order; within
 * one level they keep Version::files_ order (sorted by key, disjoint), which
 * the engine checks. */
typedef struct luda_level_file {
  luda_slice smallest_user_key;
  luda_slice largest_user_key;
  int32_t level;
  uint32_t reserved;
} luda_level_file;

/* One compaction (leveldb::Compaction, db/version_set.h:321-391). */
typedef struct luda_job {
  int32_t level;              /* Compaction::level() */
  uint32_t flags;             /* LUDA_JOB_* */
  uint64_t smallest_snapshot; /* CompactionState::smallest_snapshot, db/db_impl.cc:1073-1077 */
  luda_options options;
  const luda_file* inputs[2]; /* Compaction::inputs_[0], inputs_[1], in Compaction order */
  uint32_t num_inputs[2];
  const luda_grandparent* grandparents;
  uint32_t num_grandparents;
  uint32_t num_deeper;
  const luda_level_file* deeper;
} luda_job;

/* One output table: CompactionState::Output (db/db_impl.cc:195-227) without
 * the file number, which the caller assigns in output order. */
typedef struct luda_output {
  uint64_t offset;      /* byte offset of this table inside result.data / result.device_data */
  uint64_t file_size;   /* Output::file_size (TableBuilder::FileSize after Finish) */
  uint64_t num_entries; /* TableBuilder::NumEntries */
  uint64_t smallest_offset; /* Output::smallest, an internal key in result.keys */
  uint64_t largest_offset;  /* Output::largest */
  uint32_t smallest_size;
  uint32_t largest_size;
} luda_output;

/* Timings and byte counts of one call (CompactionStats, db/db_impl.h:92-104). */
typedef struct luda_stats {
  double device_ms;         /* inputs resident in HBM -> outputs resident in HBM */
  double total_ms;          /* whole call, host to host */
  double h2d_ms;            /* input upload, 0 when inputs are already on the device */
  double d2h_ms;            /* output download, 0 with LUDA_JOB_KEEP_ON_DEVICE */
  double top_kernel_ms;     /* summed duration of the most expensive kernel (LUDA_KERNEL_TIMING set) */
  uint64_t top_kernel_launches;
  char top_kernel[48];      /* its name */
  uint64_t bytes_read;      /* sum of input file sizes */
  uint64_t bytes_written;   /* sum of output file sizes */
  uint64_t entries_in;
  uint64_t entries_kept;
  uint64_t blocks_in;
  uint64_t blocks_out;
  uint32_t attempts;        /* 1 unless a capacity guess had to be raised */
  uint32_t key_width;       /* user-key bytes held per entry in the key planes */
  uint64_t data_block_bytes;  /* output data blocks with their trailers */
  uint64_t top_kernel_bytes;  /* algorithmic HBM bytes of top_kernel's launches in this call */
  double end_clock_ms;        /* CLOCK_MONOTONIC, in milliseconds, when the result was complete */
  uint32_t decode_ahead;      /* 1 when the decoder launched behind the input scan, before the host had
                                 read its counts (sized like the context's last jobs), fitted them and
                                 was used; 0 when the decoder ran after the counts were known */
  uint32_t reserved_stats;
} luda_stats;

typedef struct luda_result {
  uint32_t num_outputs;
  uint32_t reserved;
  luda_output* outputs;
  uint8_t* keys;          /* smallest/largest key bytes */
  uint8_t* data;          /* host copy of all tables, concatenated; NULL with KEEP_ON_DEVICE */
  void* device_data;      /* the same bytes in HBM, owned by the result */
  uint64_t total_bytes;
  luda_stats stats;
} luda_result;

typedef struct luda_ctx luda_ctx;

/* Creates an engine on HIP device `device`.  hbm_budget_bytes bounds the
 * device memory one call may hold (0 = 64 GiB).  Replaces the per-DB
 * gpu::HostAndDeviceMemory (cuda/decode_kv.cu:43-168, db/db_impl.h:215).
 * Jobs that should run side by side on one device go through one context's
 * worker slots (luda_submit): its single-pass scans are built for that.  Two
 * contexts compacting on one device at once can stall each other's scans,
 * which then fail the job (LUDA_NOT_SUPPORTED, "prefix sum") after about two
 * seconds rather than hang. */
int luda_init(int device, uint64_t hbm_budget_bytes, luda_ctx** out, char* err, size_t errlen);
/* Frees the context and its device memory: results' device tables live in
 * it, so free every result first.  The context keeps device memory it has
 * used (working buffers, freed outputs) for later calls, giving a wholly
 * free 1 GiB-or-larger piece back only when 8 GiB more than is in use would
 * still be held. */
void luda_shutdown(luda_ctx* ctx);

/* Runs one compaction.  Replaces db/db_impl.cc:1079-1197 (the merge, drop,
 * cut and build loop of DoCompactionWork).  On LUDA_OK *out holds the output
 * tables in the order DoCompactionWork would have opened them. */
int luda_compact(luda_ctx* ctx, const luda_job* job, luda_result** out, char* err, size_t errlen);

/* Builds tables from an already sorted stream of internal keys and values,
 * with no dropping: the memtable flush of db/builder.cc:17-77 when
 * opts->max_file_size is UINT64_MAX (one table), otherwise cut into tables at
 * the point DoCompactionWork would cut (db/db_impl.cc:1173-1180).
 * Entry i is keys[key_off[i] .. key_off[i+1]) and vals[val_off[i] .. val_off[i+1]).
 * `flags` takes LUDA_JOB_KEEP_ON_DEVICE. */
int luda_build_tables(luda_ctx* ctx, const luda_options* opts, uint64_t n, const uint8_t* keys,
                      const uint64_t* key_off, const uint8_t* vals, const uint64_t* val_off,
                      uint32_t flags, luda_result** out, char* err, size_t errlen);

void luda_free_result(luda_ctx* ctx, luda_result* res);
/* Verifies tables the way Table::Open plus a full read with
 * ReadOptions::verify_checksums does (table/table.cc:38-134, ReadBlock
 * table/format.cc:83-96, which the fork leaves unchecked at :88): the footer,
 * the index block, the metaindex block and the blocks it names (the filter),
 * and every data block's masked CRC-32C over the block and its type byte.
 * LUDA_OK when every table passes; LUDA_CORRUPTION otherwise, with
 * *first_bad the lowest failing table when a checksum or the footer/metaindex
 * check failed (UINT32_MAX when the fault was found while the index blocks
 * were read).  `flags` takes LUDA_JOB_INPUTS_ON_DEVICE.  A caller can run it
 * on a compaction's outputs where DBImpl::FinishCompactionOutputFile
 * re-opens them (db/db_impl.cc:1023-1033). */
int luda_verify_tables(luda_ctx* ctx, const luda_file* tables, uint32_t n, uint32_t flags, uint32_t* first_bad,
                       char* err, size_t errlen);

/* luda_verify_tables for callers that cannot allocate memory (code under
 * HIP graph capture, serving loops that forbid allocation).  It has no
 * counterpart in the reference.  The call allocates nothing: its working
 * memory is a fixed array in the library's code object, one per device (72
 * MiB, held until the process ends), and the staging buffers luda_init
 * reserved.  Tables must be device-resident (flags must hold
 * LUDA_JOB_INPUTS_ON_DEVICE, else LUDA_INVALID_ARGUMENT); at most 4096 tables
 * and 128 MiB of table bytes per call (LUDA_NOT_SUPPORTED beyond).  Results
 * equal luda_verify_tables'.  Every call on one device shares that device's
 * array, so calls that overlap on one device, from any context, race. */
int luda_verify_tables_noalloc(luda_ctx* ctx, const luda_file* tables, uint32_t n, uint32_t flags,
                               uint32_t* first_bad, char* err, size_t errlen);

/* ---- Reading tables ------------------------------------------------------------
 * Every entry of the given tables, in order, as Table::Open and a full
 * iteration read them (table/table.cc:38-214: the footer, the index block, a
 * TwoLevelIterator over each data block's Block::Iter, table/block.cc:
 * 224-251), decoded on the device and returned as plain host arrays.  With
 * LUDA_READ_VERIFY every block's checksum is checked first, as luda_verify_
 * tables does (the fork's ReadBlock skips it, table/format.cc:88); a failure
 * is LUDA_CORRUPTION.  `flags` also takes LUDA_JOB_INPUTS_ON_DEVICE. */
#define LUDA_READ_VERIFY 0x4u
typedef struct luda_entries {
  uint64_t n;              /* entries */
  uint32_t num_tables;
  uint32_t reserved;
  uint8_t* keys;           /* entry i's internal key: keys[key_off[i] .. key_off[i+1]) */
  uint64_t* key_off;       /* n + 1 */
  uint8_t* vals;           /* entry i's value: vals[val_off[i] .. val_off[i+1]) */
  uint64_t* val_off;       /* n + 1 */
  uint64_t* table_first;   /* num_tables + 1: first entry of each table, then n */
} luda_entries;
int luda_read_tables(luda_ctx* ctx, const luda_file* tables, uint32_t n, uint32_t flags, luda_entries** out,
                     char* err, size_t errlen);
void luda_free_entries(luda_entries* e);

/* ---- Tuning ------------------------------------------------------------------
 * Launch shapes a caller (or luda_amd/autotune.py, which measures them) can
 * set per class of job.  A rule applies to jobs whose input bytes lie in
 * [min_bytes, max_bytes) and whose longest user key, rounded up to 8 bytes,
 * lies in [min_key_bytes, max_key_bytes]; the last matching rule wins.  A
 * zero field keeps the engine's default. */
typedef struct luda_tuning {
  uint32_t write_blocks_groups; /* workgroups (of 16 waves) of the data block writer; default one per CU (256) */
  uint32_t decode_groups;       /* wavefronts of the block decoder; default min(blocks / 4, 4096) */
  uint32_t reserved[2];
} luda_tuning;
typedef struct luda_tuning_rule {
  uint64_t min_bytes;
  uint64_t max_bytes;
  uint32_t min_key_bytes;
  uint32_t max_key_bytes;
  luda_tuning tuning;
} luda_tuning_rule;
/* Replaces the context's rules (n = 0 clears them); worker slots follow. */
int luda_set_tuning(luda_ctx* ctx, const luda_tuning_rule* rules, uint32_t n);

/* Batched and served compaction: jobs queued on the context run on worker
 * slots, each with its own HIP stream and working memory, so that the
 * kernels of several jobs overlap on the device.  A queued job starts when a
 * slot is free and the device memory the running jobs are estimated to hold
 * (about 4.5x their input bytes each) leaves room for it under the context's
 * budget; a job is always admitted when no other job runs.  Replaces the
 * one-compaction-at-a-time gate of DBImpl::BackgroundCompaction
 * (db/db_impl.cc:865-871) and the GPU flush gate (:94-97); the reference has
 * no queued form.
 *
 * `done` runs on the thread that calls luda_poll or luda_drain, once per job,
 * with the job's status, its result (LUDA_OK only; free it with
 * luda_free_result on this context) and an error message.  The job, its
 * inputs and `user` must stay valid until then.  Results are the same as
 * luda_compact's for the same job. */
typedef void (*luda_done_fn)(void* user, int status, luda_result* result, const char* err);
/* Worker slots (default 6); only before the first luda_submit.  Each slot
 * runs its jobs on a HIP stream of its own; HIP maps a process's streams onto
 * GPU_MAX_HW_QUEUES hardware queues (4 unless the environment sets it before
 * the process's first HIP call), and slots whose streams share a queue run
 * their kernels in series: set it to at least slots + 2. */
int luda_set_slots(luda_ctx* ctx, uint32_t slots);
int luda_submit(luda_ctx* ctx, const luda_job* job, luda_done_fn done, void* user);
/* Queues a memtable flush: the arguments of luda_build_tables, which must
 * stay valid until `done` runs.  Flushes share the worker slots and the
 * budget with compactions, so a flush that arrives while compactions run is
 * admitted beside them.  Replaces DBImpl::CompactMemTable's BuildTable call
 * (db/db_impl.cc:645-653 with the GPU gate :94-97). */
int luda_submit_build(luda_ctx* ctx, const luda_options* opts, uint64_t n, const uint8_t* keys,
                      const uint64_t* key_off, const uint8_t* vals, const uint64_t* val_off, uint32_t flags,
                      luda_done_fn done, void* user);
/* Runs the callbacks of finished jobs; returns how many (never blocks). */
int luda_poll(luda_ctx* ctx);
/* Waits for every submitted job and runs its callback; returns how many. */
int luda_drain(luda_ctx* ctx);

/* ---- File I/O around a compaction ------------------------------------------
 * The reference reads each input file whole with fread (cuda/cuda_common.h:
 * 302-311) and writes each output on a pthread of its own with fwrite and
 * fsync (db/db_impl.cc:55-65,1490-1565).  Here input files go from disk
 * through two pinned chunk buffers into HBM on the context's I/O stream, a
 * stream of its own: reading chunk k+1 from disk overlaps the copy of chunk
 * k, and both overlap kernels a compaction on another thread is running.
 * Outputs go from HBM to pinned memory in one copy, then to their files on up
 * to 16 writer threads. */
typedef struct luda_file_set {
  uint32_t n;
  uint32_t reserved;
  luda_file* files;     /* files[i].data: device address of file i, .size, .number */
  void* device_data;    /* the one device buffer holding them */
  uint64_t total_bytes; /* sum of file sizes */
  double read_ms;       /* whole call */
} luda_file_set;
/* Reads `n` table files into HBM (numbers[i] goes to files[i].number; numbers
 * may be NULL).  The result feeds luda_compact with LUDA_JOB_INPUTS_ON_DEVICE;
 * free it with luda_free_files.  LUDA_IO_ERROR when a file cannot be read. */
int luda_read_files(luda_ctx* ctx, const char* const* paths, const uint64_t* numbers, uint32_t n,
                    luda_file_set** out, char* err, size_t errlen);
void luda_free_files(luda_ctx* ctx, luda_file_set* set);
/* Writes output table i of `res` to paths[i] (created or truncated), with
 * fsync when sync is non-zero, as FinishCompactionOutputFile's Sync/Close do
 * (db/db_impl.cc:1000-1014).  Works for results kept on the device and for
 * host results. */
int luda_write_outputs(luda_ctx* ctx, const luda_result* res, const char* const* paths, int sync, char* err,
                       size_t errlen);

/* ---- The install record ------------------------------------------------------
 * A VersionEdit as VersionEdit::EncodeTo serialises it (db/version_edit.cc:
 * 41-89), and the MANIFEST log record that carries it (log::Writer::AddRecord,
 * db/log_writer.cc:35-105).  Host-only; no context needed. */
#define LUDA_EDIT_HAS_COMPARATOR 0x1u
#define LUDA_EDIT_HAS_LOG_NUMBER 0x2u
#define LUDA_EDIT_HAS_PREV_LOG_NUMBER 0x4u
#define LUDA_EDIT_HAS_NEXT_FILE 0x8u
#define LUDA_EDIT_HAS_LAST_SEQUENCE 0x10u
typedef struct luda_edit_pointer {
  int32_t level;
  uint32_t reserved;
  luda_slice key; /* an internal key */
} luda_edit_pointer;
typedef struct luda_edit_deleted {
  int32_t level;
  uint32_t reserved;
  uint64_t number;
} luda_edit_deleted;
typedef struct luda_edit_file {
  int32_t level;
  uint32_t reserved;
  uint64_t number;
  uint64_t file_size;
  luda_slice smallest; /* internal keys */
  luda_slice largest;
} luda_edit_file;
/* VersionEdit's fields (db/version_edit.h:86-100): the deleted files are a
 * set (sorted by level, then number; duplicates once), the others keep their
 * order. */
typedef struct luda_edit {
  uint32_t has; /* LUDA_EDIT_HAS_* */
  uint32_t num_compact_pointers;
  luda_slice comparator;
  uint64_t log_number;
  uint64_t prev_log_number;
  uint64_t next_file_number;
  uint64_t last_sequence;
  const luda_edit_pointer* compact_pointers;
  const luda_edit_deleted* deleted;
  uint32_t num_deleted;
  uint32_t num_new_files;
  const luda_edit_file* new_files;
} luda_edit;
/* Serialises e into dst[0 .. *len).  If cap is too small, returns
 * LUDA_INVALID_ARGUMENT with *len = the bytes needed. */
int luda_encode_edit(const luda_edit* e, uint8_t* dst, uint64_t cap, uint64_t* len);
/* The edit DBImpl::InstallCompactionResults applies (db/db_impl.cc:1040-1056):
 * every input file deleted from its level (Compaction::AddInputDeletions,
 * db/version_set.cc:1527-1533), then output i added at level + 1 as file
 * output_numbers[i] with its size and boundary keys.  `base` (may be NULL)
 * supplies the header fields and compact pointers LogAndApply and
 * SetupOtherInputs put into the same edit. */
int luda_encode_compaction_edit(const luda_job* job, const luda_result* res, const uint64_t* output_numbers,
                                const luda_edit* base, uint8_t* dst, uint64_t cap, uint64_t* len);
/* Appends `payload` as one log record to a log whose current block offset is
 * *block_offset (the file length mod 32768), updating it: the record split
 * into FULL/FIRST/MIDDLE/LAST fragments with 7-byte headers (masked CRC-32C
 * of type and data, length, type) and block trailers zero-filled.  Needs at
 * most n + 7 * (n / 32761 + 2) + 6 bytes; returns LUDA_INVALID_ARGUMENT with
 * *len = the bytes needed if cap is too small. */
int luda_log_record(uint64_t* block_offset, const uint8_t* payload, uint64_t n, uint8_t* dst, uint64_t cap,
                    uint64_t* len);

/* Device-buffer helpers so that a caller with no HIP code of its own can keep
 * inputs resident in HBM (LUDA_JOB_INPUTS_ON_DEVICE) and read outputs back. */
int luda_device_alloc(luda_ctx* ctx, uint64_t bytes, void** dptr);
int luda_device_free(luda_ctx* ctx, void* dptr);
int luda_memcpy_h2d(luda_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int luda_memcpy_d2h(luda_ctx* ctx, void* dst, const void* src, uint64_t bytes);

const char* luda_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LUDA_H_ */

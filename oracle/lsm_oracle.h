/*
 * lsm_oracle.h — CPU restatement of the xpsair/luda LevelDB fork's table
 * writer, table reader and compaction loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP engine
 * is compared against, and the "port" CPU baseline that the benchmark times.
 * Only tests/, __graft_entry__.smoke() and the benchmark's CPU-baseline leg
 * may load it; the engine under luda_amd/ never does.
 *
 * Every function cites the reference file:line it restates.  The job and
 * result structs are the engine's own (include/luda.h) so that the tests can
 * hand the same job to the engine, to this oracle and to the compiled
 * reference under oracle/_ref.
 */
#ifndef LSM_ORACLE_H_
#define LSM_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#include "../include/luda.h"

#ifdef __cplusplus
extern "C" {
#endif

/* util/coding.cc */
size_t or_encode_varint64(uint8_t* dst, uint64_t v);
size_t or_varint_length(uint64_t v);
/* util/crc32c.{h,cc} */
uint32_t or_crc32c_extend(uint32_t crc, const uint8_t* data, size_t n);
uint32_t or_crc32c_value(const uint8_t* data, size_t n);
uint32_t or_crc32c_mask(uint32_t crc);
uint32_t or_crc32c_unmask(uint32_t masked);
/* util/hash.cc */
uint32_t or_hash(const uint8_t* data, size_t n, uint32_t seed);
/* util/bloom.cc: writes the filter for n keys (key i = keys[off[i]..off[i+1]))
 * into dst; returns its length, or 0 if cap is too small. */
size_t or_bloom_create(const uint8_t* keys, const uint64_t* off, uint32_t n, uint32_t bits_per_key,
                       uint8_t* dst, size_t cap);
int or_bloom_may_match(const uint8_t* key, size_t klen, const uint8_t* filter, size_t flen);
/* db/dbformat.cc */
int or_ikey_compare(const uint8_t* a, size_t alen, const uint8_t* b, size_t blen);
size_t or_shortest_separator(const uint8_t* start, size_t slen, const uint8_t* limit, size_t llen,
                             uint8_t* out);
size_t or_short_successor(const uint8_t* key, size_t klen, uint8_t* out);

/* table/filter_block.cc, with a choice of policy: 0 = the bloom policy over
 * user keys (InternalFilterPolicy + BloomFilterPolicy), 1 = the test policy of
 * table/filter_block_test.cc:16-36 over whole keys. */
typedef struct or_filter_builder or_filter_builder;
or_filter_builder* or_filter_new(int policy, uint32_t bits_per_key);
void or_filter_start_block(or_filter_builder* f, uint64_t block_offset);
void or_filter_add_key(or_filter_builder* f, const uint8_t* key, size_t n);
/* Finishes the block; returns a pointer valid until or_filter_free. */
const uint8_t* or_filter_finish(or_filter_builder* f, size_t* len);
void or_filter_free(or_filter_builder* f);
int or_filter_key_may_match(int policy, const uint8_t* block, size_t blen, uint64_t block_offset,
                            const uint8_t* key, size_t klen);

/* Table reader (table/table.cc:38-214, table/block.cc, table/format.cc) with
 * the block CRC check restored.  Calls fn for every entry in order.  Returns
 * LUDA_OK or LUDA_CORRUPTION; *nentries gets the count. */
typedef void (*or_entry_fn)(void* user, const uint8_t* key, size_t klen, const uint8_t* val,
                            size_t vlen);
int or_table_iterate(const uint8_t* data, uint64_t size, int verify_crc, or_entry_fn fn, void* user,
                     uint64_t* nentries);

/* db/builder.cc:17-77 / TableBuilder cut at max_file_size like
 * db/db_impl.cc:1173-1180.  Same contract as luda_build_tables. */
int or_build_tables(const luda_options* opts, uint64_t n, const uint8_t* keys, const uint64_t* key_off,
                    const uint8_t* vals, const uint64_t* val_off, luda_result** out);

/* db/db_impl.cc:1060-1224 (CPU DoCompactionWork).  Input data pointers are
 * host pointers. */
int or_compact(const luda_job* job, luda_result** out);

void or_free_result(luda_result* r);

#ifdef __cplusplus
}
#endif

#endif /* LSM_ORACLE_H_ */

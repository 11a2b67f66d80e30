/*
 * lsm_oracle.c — CPU restatement of the xpsair/luda LevelDB fork's table
 * format and CPU compaction.  TEST INFRASTRUCTURE ONLY: this is the checker
 * the HIP engine is compared against and the "port" CPU baseline; the engine
 * never links or calls it.
 *
 * Pinning (what checks each part, see tests/test_oracle*.py):
 *   coding, crc32c, hash      known answers of util/coding_test.cc:56-76,
 *                             util/crc32c_test.cc:13-54, util/hash_test.cc:12-40
 *   separators / successor    db/dbformat_test.cc:46-107 expected outputs
 *   filter block framing      table/filter_block_test.cc:44-120
 *   bloom                     closed form + util/bloom_test.cc behaviour
 *   whole tables, compaction  byte equality with the reference's own
 *                             TableBuilder / MergingIterator / Compaction code
 *                             compiled from its sources (oracle/_ref), on
 *                             seeded workloads, plus brute-force invariants
 */
#include "lsm_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* growable byte buffer                                                       */

typedef struct {
  uint8_t* p;
  size_t n, cap;
} buf_t;

static void buf_reserve(buf_t* b, size_t need) {
  if (need <= b->cap) return;
  size_t c = b->cap ? b->cap : 64;
  while (c < need) c *= 2;
  uint8_t* q = (uint8_t*)realloc(b->p, c);
  if (!q) {
    fprintf(stderr, "lsm_oracle: out of memory\n");
    abort();
  }
  b->p = q;
  b->cap = c;
}
static void buf_append(buf_t* b, const void* d, size_t n) {
  buf_reserve(b, b->n + n);
  if (n) memcpy(b->p + b->n, d, n);
  b->n += n;
}
static void buf_push(buf_t* b, uint8_t c) { buf_append(b, &c, 1); }
static void buf_resize(buf_t* b, size_t n) {
  buf_reserve(b, n);
  if (n > b->n) memset(b->p + b->n, 0, n - b->n);
  b->n = n;
}
static void buf_assign(buf_t* b, const void* d, size_t n) {
  b->n = 0;
  buf_append(b, d, n);
}
static void buf_free(buf_t* b) {
  free(b->p);
  b->p = NULL;
  b->n = b->cap = 0;
}

/* ------------------------------------------------------------------------ */
/* util/coding.cc:9-110                                                       */

static void put_fixed32(buf_t* b, uint32_t v) {
  uint8_t t[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
  buf_append(b, t, 4);
}
static void put_fixed64(buf_t* b, uint64_t v) {
  put_fixed32(b, (uint32_t)v);
  put_fixed32(b, (uint32_t)(v >> 32));
}
size_t or_encode_varint64(uint8_t* dst, uint64_t v) {
  size_t i = 0;
  while (v >= 128) {
    dst[i++] = (uint8_t)(v | 128);
    v >>= 7;
  }
  dst[i++] = (uint8_t)v;
  return i;
}
size_t or_varint_length(uint64_t v) {
  size_t len = 1;
  while (v >= 128) {
    v >>= 7;
    len++;
  }
  return len;
}
/* EncodeVarint32 and EncodeVarint64 give the same bytes for values < 2^32. */
static void put_varint(buf_t* b, uint64_t v) {
  uint8_t t[10];
  buf_append(b, t, or_encode_varint64(t, v));
}
static uint32_t dec32(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static uint64_t dec64(const uint8_t* p) { return (uint64_t)dec32(p) | ((uint64_t)dec32(p + 4) << 32); }

/* util/coding.cc:121-140 GetVarint32PtrFallback */
static const uint8_t* get_varint32(const uint8_t* p, const uint8_t* limit, uint32_t* v) {
  uint32_t result = 0;
  for (uint32_t shift = 0; shift <= 28 && p < limit; shift += 7) {
    uint32_t byte = *p++;
    if (byte & 128) {
      result |= (byte & 127) << shift;
    } else {
      result |= byte << shift;
      *v = result;
      return p;
    }
  }
  return NULL;
}
/* util/coding.cc:150-165 GetVarint64Ptr */
static const uint8_t* get_varint64(const uint8_t* p, const uint8_t* limit, uint64_t* v) {
  uint64_t result = 0;
  for (uint32_t shift = 0; shift <= 63 && p < limit; shift += 7) {
    uint64_t byte = *p++;
    if (byte & 128) {
      result |= (byte & 127) << shift;
    } else {
      result |= byte << shift;
      *v = result;
      return p;
    }
  }
  return NULL;
}

/* ------------------------------------------------------------------------ */
/* util/crc32c.cc:276-377 (the function; the table is derived from the       */
/* Castagnoli polynomial, reflected 0x82F63B78, not copied)                   */

static uint32_t g_crc_table[256];
static int g_crc_ready = 0;

static void crc_init(void) {
  if (g_crc_ready) return;
  for (uint32_t i = 0; i < 256; i++) {
    uint32_t c = i;
    for (int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ 0x82F63B78u : (c >> 1);
    g_crc_table[i] = c;
  }
  g_crc_ready = 1;
}
uint32_t or_crc32c_extend(uint32_t crc, const uint8_t* p, size_t n) {
  crc_init();
  uint32_t l = crc ^ 0xffffffffu;
  for (size_t i = 0; i < n; i++) l = g_crc_table[(l ^ p[i]) & 0xff] ^ (l >> 8);
  return l ^ 0xffffffffu;
}
uint32_t or_crc32c_value(const uint8_t* p, size_t n) { return or_crc32c_extend(0, p, n); }
/* util/crc32c.h:29-38 */
uint32_t or_crc32c_mask(uint32_t crc) { return ((crc >> 15) | (crc << 17)) + 0xa282ead8u; }
uint32_t or_crc32c_unmask(uint32_t m) {
  uint32_t rot = m - 0xa282ead8u;
  return (rot >> 17) | (rot << 15);
}

/* ------------------------------------------------------------------------ */
/* util/hash.cc:22-53                                                         */

uint32_t or_hash(const uint8_t* data, size_t n, uint32_t seed) {
  const uint32_t m = 0xc6a4a793u;
  const uint32_t r = 24;
  const uint8_t* limit = data + n;
  uint32_t h = seed ^ (uint32_t)(n * m);
  while (data + 4 <= limit) {
    uint32_t w = dec32(data);
    data += 4;
    h += w;
    h *= m;
    h ^= (h >> 16);
  }
  switch (limit - data) {
    case 3:
      h += (uint32_t)data[2] << 16;
      /* fall through */
    case 2:
      h += (uint32_t)data[1] << 8;
      /* fall through */
    case 1:
      h += data[0];
      h *= m;
      h ^= (h >> r);
      break;
  }
  return h;
}

/* ------------------------------------------------------------------------ */
/* util/bloom.cc:17-85                                                        */

typedef struct {
  const uint8_t* p;
  size_t n;
} slice_t;

static size_t bloom_k(uint32_t bits_per_key) {
  size_t k = (size_t)(bits_per_key * 0.69); /* util/bloom.cc:21 */
  if (k < 1) k = 1;
  if (k > 30) k = 30;
  return k;
}
/* BloomFilterPolicy::CreateFilter, util/bloom.cc:28-54 */
static void bloom_append(buf_t* dst, const slice_t* keys, size_t n, uint32_t bits_per_key) {
  size_t k = bloom_k(bits_per_key);
  size_t bits = n * (size_t)bits_per_key;
  if (bits < 64) bits = 64;
  size_t bytes = (bits + 7) / 8;
  bits = bytes * 8;
  size_t init = dst->n;
  buf_resize(dst, init + bytes);
  buf_push(dst, (uint8_t)k);
  uint8_t* array = dst->p + init;
  for (size_t i = 0; i < n; i++) {
    uint32_t h = or_hash(keys[i].p, keys[i].n, 0xbc9f1d34u);
    const uint32_t delta = (h >> 17) | (h << 15);
    for (size_t j = 0; j < k; j++) {
      const uint32_t bitpos = (uint32_t)(h % bits);
      array[bitpos / 8] |= (uint8_t)(1 << (bitpos % 8));
      h += delta;
    }
  }
}
size_t or_bloom_create(const uint8_t* keys, const uint64_t* off, uint32_t n, uint32_t bits_per_key,
                       uint8_t* dst, size_t cap) {
  slice_t* s = (slice_t*)malloc(sizeof(slice_t) * (n ? n : 1));
  for (uint32_t i = 0; i < n; i++) {
    s[i].p = keys + off[i];
    s[i].n = off[i + 1] - off[i];
  }
  buf_t b = {0};
  bloom_append(&b, s, n, bits_per_key);
  free(s);
  size_t len = b.n;
  if (len > cap) len = 0;
  else memcpy(dst, b.p, b.n);
  buf_free(&b);
  return len;
}
/* BloomFilterPolicy::KeyMayMatch, util/bloom.cc:56-79 */
int or_bloom_may_match(const uint8_t* key, size_t klen, const uint8_t* array, size_t len) {
  if (len < 2) return 0;
  const size_t bits = (len - 1) * 8;
  const size_t k = array[len - 1];
  if (k > 30) return 1;
  uint32_t h = or_hash(key, klen, 0xbc9f1d34u);
  const uint32_t delta = (h >> 17) | (h << 15);
  for (size_t j = 0; j < k; j++) {
    const uint32_t bitpos = (uint32_t)(h % bits);
    if ((array[bitpos / 8] & (1 << (bitpos % 8))) == 0) return 0;
    h += delta;
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* db/dbformat.{h,cc}, util/comparator.cc                                     */

static int bytewise_compare(const uint8_t* a, size_t an, const uint8_t* b, size_t bn) {
  size_t m = an < bn ? an : bn;
  int r = m ? memcmp(a, b, m) : 0;
  if (r == 0) r = an < bn ? -1 : (an > bn ? 1 : 0);
  return r < 0 ? -1 : (r > 0 ? 1 : 0);
}
/* InternalKeyComparator::Compare, db/dbformat.cc:51-67 */
int or_ikey_compare(const uint8_t* a, size_t alen, const uint8_t* b, size_t blen) {
  int r = bytewise_compare(a, alen - 8, b, blen - 8);
  if (r == 0) {
    uint64_t an = dec64(a + alen - 8), bn = dec64(b + blen - 8);
    if (an > bn) r = -1;
    else if (an < bn) r = 1;
  }
  return r;
}
/* (kMaxSequenceNumber << 8) | kValueTypeForSeek, db/dbformat.cc:14-18, dbformat.h:59,65 */
static const uint64_t kMaxSeqTag = ((((uint64_t)1 << 56) - 1) << 8) | 1;

/* InternalKeyComparator::FindShortestSeparator, db/dbformat.cc:69-86, with
 * BytewiseComparatorImpl::FindShortestSeparator, util/comparator.cc:29-48 */
static void separator(buf_t* start, const uint8_t* limit, size_t llen) {
  const uint8_t* us = start->p;
  size_t usn = start->n - 8;
  const uint8_t* ul = limit;
  size_t uln = llen - 8;
  buf_t tmp = {0};
  buf_assign(&tmp, us, usn);
  size_t min_length = usn < uln ? usn : uln;
  size_t diff = 0;
  while (diff < min_length && tmp.p[diff] == ul[diff]) diff++;
  if (diff < min_length) {
    uint8_t db = tmp.p[diff];
    if (db < 0xff && db + 1 < ul[diff]) {
      tmp.p[diff]++;
      tmp.n = diff + 1;
    }
  }
  if (tmp.n < usn && bytewise_compare(us, usn, tmp.p, tmp.n) < 0) {
    put_fixed64(&tmp, kMaxSeqTag);
    buf_assign(start, tmp.p, tmp.n);
  }
  buf_free(&tmp);
}
/* InternalKeyComparator::FindShortSuccessor, db/dbformat.cc:88-101, with
 * BytewiseComparatorImpl::FindShortSuccessor, util/comparator.cc:50-63 */
static void successor(buf_t* key) {
  size_t un = key->n - 8;
  buf_t tmp = {0};
  buf_assign(&tmp, key->p, un);
  for (size_t i = 0; i < un; i++) {
    if (tmp.p[i] != 0xff) {
      tmp.p[i]++;
      tmp.n = i + 1;
      break;
    }
  }
  if (tmp.n < un && bytewise_compare(key->p, un, tmp.p, tmp.n) < 0) {
    put_fixed64(&tmp, kMaxSeqTag);
    buf_assign(key, tmp.p, tmp.n);
  }
  buf_free(&tmp);
}
size_t or_shortest_separator(const uint8_t* start, size_t slen, const uint8_t* limit, size_t llen,
                             uint8_t* out) {
  buf_t s = {0};
  buf_assign(&s, start, slen);
  separator(&s, limit, llen);
  size_t n = s.n;
  memcpy(out, s.p, n);
  buf_free(&s);
  return n;
}
size_t or_short_successor(const uint8_t* key, size_t klen, uint8_t* out) {
  buf_t s = {0};
  buf_assign(&s, key, klen);
  successor(&s);
  size_t n = s.n;
  memcpy(out, s.p, n);
  buf_free(&s);
  return n;
}

/* ------------------------------------------------------------------------ */
/* table/block_builder.cc:41-114 — the fork's format: each restart word is    */
/* (offset << 8) | entries_in_group                                            */

typedef struct {
  buf_t buf;
  uint32_t* restarts;
  size_t nrest, caprest;
  uint32_t counter;
  uint32_t interval;
  buf_t last_key;
} bb_t;

static void bb_push_restart(bb_t* b, uint32_t v) {
  if (b->nrest == b->caprest) {
    b->caprest = b->caprest ? b->caprest * 2 : 16;
    b->restarts = (uint32_t*)realloc(b->restarts, b->caprest * sizeof(uint32_t));
  }
  b->restarts[b->nrest++] = v;
}
static void bb_init(bb_t* b, uint32_t interval) {
  memset(b, 0, sizeof(*b));
  b->interval = interval;
  bb_push_restart(b, 0);
}
static void bb_reset(bb_t* b) {
  b->buf.n = 0;
  b->nrest = 0;
  bb_push_restart(b, 0);
  b->counter = 0;
  b->last_key.n = 0;
}
static void bb_free(bb_t* b) {
  buf_free(&b->buf);
  buf_free(&b->last_key);
  free(b->restarts);
}
static size_t bb_estimate(const bb_t* b) { return b->buf.n + b->nrest * 4 + 4; }
static int bb_empty(const bb_t* b) { return b->buf.n == 0; }
static void bb_finish(bb_t* b) {
  if (b->nrest) b->restarts[b->nrest - 1] = (b->restarts[b->nrest - 1] << 8) | b->counter;
  for (size_t i = 0; i < b->nrest; i++) put_fixed32(&b->buf, b->restarts[i]);
  put_fixed32(&b->buf, (uint32_t)b->nrest);
}
static void bb_add(bb_t* b, const uint8_t* key, size_t klen, const uint8_t* val, size_t vlen) {
  size_t shared = 0;
  if (b->counter < b->interval) {
    size_t m = b->last_key.n < klen ? b->last_key.n : klen;
    while (shared < m && b->last_key.p[shared] == key[shared]) shared++;
  } else {
    bb_push_restart(b, (uint32_t)b->buf.n);
    b->restarts[b->nrest - 2] = (b->restarts[b->nrest - 2] << 8) | b->counter;
    b->counter = 0;
  }
  size_t non_shared = klen - shared;
  put_varint(&b->buf, shared);
  put_varint(&b->buf, non_shared);
  put_varint(&b->buf, vlen);
  buf_append(&b->buf, key + shared, non_shared);
  buf_append(&b->buf, val, vlen);
  b->last_key.n = shared;
  buf_append(&b->last_key, key + shared, non_shared);
  b->counter++;
}

/* ------------------------------------------------------------------------ */
/* table/filter_block.cc:14-76                                               */

struct or_filter_builder {
  int policy; /* 0: InternalFilterPolicy(bloom), 1: filter_block_test's TestHashFilter */
  uint32_t bits_per_key;
  buf_t keys;
  size_t* start;
  size_t nstart, capstart;
  buf_t result;
  uint32_t* offsets;
  size_t noff, capoff;
};

static void fb_push_start(or_filter_builder* f, size_t v) {
  if (f->nstart == f->capstart) {
    f->capstart = f->capstart ? f->capstart * 2 : 64;
    f->start = (size_t*)realloc(f->start, f->capstart * sizeof(size_t));
  }
  f->start[f->nstart++] = v;
}
static void fb_push_off(or_filter_builder* f, uint32_t v) {
  if (f->noff == f->capoff) {
    f->capoff = f->capoff ? f->capoff * 2 : 64;
    f->offsets = (uint32_t*)realloc(f->offsets, f->capoff * sizeof(uint32_t));
  }
  f->offsets[f->noff++] = v;
}
or_filter_builder* or_filter_new(int policy, uint32_t bits_per_key) {
  or_filter_builder* f = (or_filter_builder*)calloc(1, sizeof(*f));
  f->policy = policy;
  f->bits_per_key = bits_per_key;
  return f;
}
/* FilterBlockBuilder::GenerateFilter, table/filter_block.cc:52-76 */
static void fb_generate(or_filter_builder* f) {
  size_t n = f->nstart;
  if (n == 0) {
    fb_push_off(f, (uint32_t)f->result.n);
    return;
  }
  fb_push_start(f, f->keys.n);
  slice_t* ks = (slice_t*)malloc(n * sizeof(slice_t));
  for (size_t i = 0; i < n; i++) {
    ks[i].p = f->keys.p + f->start[i];
    ks[i].n = f->start[i + 1] - f->start[i];
  }
  fb_push_off(f, (uint32_t)f->result.n);
  if (f->policy == 0) {
    /* InternalFilterPolicy::CreateFilter strips the 8-byte tag, db/dbformat.cc:105-115 */
    for (size_t i = 0; i < n; i++) ks[i].n -= 8;
    bloom_append(&f->result, ks, n, f->bits_per_key);
  } else {
    for (size_t i = 0; i < n; i++) put_fixed32(&f->result, or_hash(ks[i].p, ks[i].n, 1));
  }
  free(ks);
  f->keys.n = 0;
  f->nstart = 0;
}
void or_filter_start_block(or_filter_builder* f, uint64_t block_offset) {
  uint64_t idx = block_offset / 2048;
  while (idx > f->noff) fb_generate(f);
}
void or_filter_add_key(or_filter_builder* f, const uint8_t* key, size_t n) {
  fb_push_start(f, f->keys.n);
  buf_append(&f->keys, key, n);
}
const uint8_t* or_filter_finish(or_filter_builder* f, size_t* len) {
  if (f->nstart) fb_generate(f);
  uint32_t array_offset = (uint32_t)f->result.n;
  for (size_t i = 0; i < f->noff; i++) put_fixed32(&f->result, f->offsets[i]);
  put_fixed32(&f->result, array_offset);
  buf_push(&f->result, 11);
  *len = f->result.n;
  return f->result.p;
}
void or_filter_free(or_filter_builder* f) {
  if (!f) return;
  buf_free(&f->keys);
  buf_free(&f->result);
  free(f->start);
  free(f->offsets);
  free(f);
}
/* FilterBlockReader, table/filter_block.cc:78-106.  For policy 0 `key` is the
 * user key (what InternalFilterPolicy::KeyMayMatch passes on). */
int or_filter_key_may_match(int policy, const uint8_t* data, size_t n, uint64_t block_offset,
                            const uint8_t* key, size_t klen) {
  if (n < 5) return 1;
  uint32_t base_lg = data[n - 1];
  uint32_t last_word = dec32(data + n - 5);
  if (last_word > n - 5) return 1;
  const uint8_t* offset = data + last_word;
  size_t num = (n - 5 - last_word) / 4;
  uint64_t index = block_offset >> base_lg;
  if (index < num) {
    uint32_t start = dec32(offset + index * 4);
    uint32_t limit = dec32(offset + index * 4 + 4);
    if (start <= limit && limit <= (size_t)(offset - data)) {
      const uint8_t* filter = data + start;
      size_t flen = limit - start;
      if (policy == 0) return or_bloom_may_match(key, klen, filter, flen);
      uint32_t h = or_hash(key, klen, 1);
      for (size_t i = 0; i + 4 <= flen; i += 4)
        if (h == dec32(filter + i)) return 1;
      return 0;
    } else if (start == limit) {
      return 0;
    }
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* table/table_builder.cc:21-264                                             */

typedef struct {
  luda_options o;
  buf_t* out;
  uint64_t base; /* where this table starts inside out */
  uint64_t offset;
  bb_t data, index;
  buf_t last_key;
  int64_t num_entries;
  or_filter_builder* filter;
  int pending_index_entry;
  uint64_t ph_off, ph_size;
} tb_t;

static void tb_init(tb_t* t, const luda_options* o, buf_t* out) {
  memset(t, 0, sizeof(*t));
  t->o = *o;
  t->out = out;
  t->base = out->n;
  bb_init(&t->data, o->block_restart_interval);
  bb_init(&t->index, 1); /* table_builder.cc:31 */
  if (o->bloom_bits_per_key) {
    t->filter = or_filter_new(0, o->bloom_bits_per_key);
    or_filter_start_block(t->filter, 0);
  }
}
static void tb_free(tb_t* t) {
  bb_free(&t->data);
  bb_free(&t->index);
  buf_free(&t->last_key);
  or_filter_free(t->filter);
}
static void handle_encode(buf_t* b, uint64_t off, uint64_t size) {
  put_varint(b, off);
  put_varint(b, size);
}
/* TableBuilder::WriteRawBlock, table_builder.cc:178-195 */
static void tb_write_raw(tb_t* t, const uint8_t* contents, size_t n, uint8_t type, uint64_t* hoff,
                         uint64_t* hsize) {
  *hoff = t->offset;
  *hsize = n;
  buf_append(t->out, contents, n);
  uint8_t trailer[5];
  trailer[0] = type;
  uint32_t crc = or_crc32c_value(contents, n);
  crc = or_crc32c_extend(crc, trailer, 1);
  uint32_t m = or_crc32c_mask(crc);
  trailer[1] = (uint8_t)m;
  trailer[2] = (uint8_t)(m >> 8);
  trailer[3] = (uint8_t)(m >> 16);
  trailer[4] = (uint8_t)(m >> 24);
  buf_append(t->out, trailer, 5);
  t->offset += n + 5;
}
/* TableBuilder::WriteBlock, table_builder.cc:142-176 (kNoCompression only) */
static void tb_write_block(tb_t* t, bb_t* b, uint64_t* hoff, uint64_t* hsize) {
  bb_finish(b);
  tb_write_raw(t, b->buf.p, b->buf.n, 0, hoff, hsize);
  bb_reset(b);
}
/* TableBuilder::Flush, table_builder.cc:126-140 */
static void tb_flush(tb_t* t) {
  if (bb_empty(&t->data)) return;
  tb_write_block(t, &t->data, &t->ph_off, &t->ph_size);
  t->pending_index_entry = 1;
  if (t->filter) or_filter_start_block(t->filter, t->offset);
}
/* TableBuilder::Add, table_builder.cc:94-124 */
static void tb_add(tb_t* t, const uint8_t* key, size_t klen, const uint8_t* val, size_t vlen) {
  if (t->pending_index_entry) {
    separator(&t->last_key, key, klen);
    buf_t h = {0};
    handle_encode(&h, t->ph_off, t->ph_size);
    bb_add(&t->index, t->last_key.p, t->last_key.n, h.p, h.n);
    buf_free(&h);
    t->pending_index_entry = 0;
  }
  if (t->filter) or_filter_add_key(t->filter, key, klen);
  buf_assign(&t->last_key, key, klen);
  t->num_entries++;
  bb_add(&t->data, key, klen, val, vlen);
  if (bb_estimate(&t->data) >= t->o.block_size) tb_flush(t);
}
/* TableBuilder::Finish, table_builder.cc:199-254 */
static void tb_finish(tb_t* t) {
  tb_flush(t);
  uint64_t foff = 0, fsize = 0, moff, msize, ioff, isize;
  if (t->filter) {
    size_t flen;
    const uint8_t* fb = or_filter_finish(t->filter, &flen);
    tb_write_raw(t, fb, flen, 0, &foff, &fsize);
  }
  {
    bb_t meta;
    bb_init(&meta, t->o.block_restart_interval);
    if (t->filter) {
      static const char kName[] = "filter.leveldb.BuiltinBloomFilter2";
      buf_t h = {0};
      handle_encode(&h, foff, fsize);
      bb_add(&meta, (const uint8_t*)kName, sizeof(kName) - 1, h.p, h.n);
      buf_free(&h);
    }
    tb_write_block(t, &meta, &moff, &msize);
    bb_free(&meta);
  }
  if (t->pending_index_entry) {
    successor(&t->last_key);
    buf_t h = {0};
    handle_encode(&h, t->ph_off, t->ph_size);
    bb_add(&t->index, t->last_key.p, t->last_key.n, h.p, h.n);
    buf_free(&h);
    t->pending_index_entry = 0;
  }
  tb_write_block(t, &t->index, &ioff, &isize);
  /* Footer::EncodeTo, table/format.cc:31-40 */
  buf_t f = {0};
  handle_encode(&f, moff, msize);
  handle_encode(&f, ioff, isize);
  buf_resize(&f, 40);
  put_fixed32(&f, 0x8b80fb57u);
  put_fixed32(&f, 0xdb477524u);
  buf_append(t->out, f.p, f.n);
  t->offset += f.n;
  buf_free(&f);
}

/* ------------------------------------------------------------------------ */
/* results                                                                    */

typedef struct {
  luda_output* outs;
  size_t n, cap;
  buf_t keys;
  buf_t data;
} results_t;

static void res_add(results_t* r, uint64_t offset, uint64_t size, uint64_t nent, const buf_t* smallest,
                    const buf_t* largest) {
  if (r->n == r->cap) {
    r->cap = r->cap ? r->cap * 2 : 16;
    r->outs = (luda_output*)realloc(r->outs, r->cap * sizeof(luda_output));
  }
  luda_output* o = &r->outs[r->n++];
  memset(o, 0, sizeof(*o));
  o->offset = offset;
  o->file_size = size;
  o->num_entries = nent;
  o->smallest_offset = r->keys.n;
  o->smallest_size = (uint32_t)smallest->n;
  buf_append(&r->keys, smallest->p, smallest->n);
  o->largest_offset = r->keys.n;
  o->largest_size = (uint32_t)largest->n;
  buf_append(&r->keys, largest->p, largest->n);
}
static luda_result* res_finish(results_t* r, uint64_t bytes_read) {
  luda_result* out = (luda_result*)calloc(1, sizeof(luda_result));
  out->num_outputs = (uint32_t)r->n;
  out->outputs = r->outs;
  out->keys = r->keys.p;
  out->data = r->data.p;
  out->total_bytes = r->data.n;
  out->stats.bytes_read = bytes_read;
  out->stats.bytes_written = r->data.n;
  out->stats.attempts = 1;
  return out;
}
void or_free_result(luda_result* r) {
  if (!r) return;
  free(r->outputs);
  free(r->keys);
  free(r->data);
  free(r);
}

/* ------------------------------------------------------------------------ */
/* table reader: Table::Open/NewIterator (table/table.cc:38-214), ReadBlock    */
/* (table/format.cc:65-154) with the CRC check restored, Block::Iter           */
/* (table/block.cc:24-251)                                                     */

typedef struct {
  const uint8_t* p;
  uint64_t n;
} span_t;

static int read_block(const uint8_t* data, uint64_t size, uint64_t off, uint64_t n, int verify,
                      span_t* out) {
  if (off > size || n > size || off + n + 5 > size) return LUDA_CORRUPTION;
  const uint8_t* b = data + off;
  if (verify) {
    uint32_t crc = or_crc32c_unmask(dec32(b + n + 1));
    if (or_crc32c_value(b, n + 1) != crc) return LUDA_CORRUPTION;
  }
  /* kSnappyCompression cannot be decoded here; every other type byte is
   * taken as raw, as the fork's ReadBlock does (table/format.cc:134-150). */
  if (b[n] == 1) return LUDA_NOT_SUPPORTED;
  out->p = b;
  out->n = n;
  return LUDA_OK;
}

/* Iterates one block; calls fn(key, value).  Block::Block size checks at
 * table/block.cc:24-38, DecodeEntry at :54-74, ParseNextKey at :224-251. */
static int block_iterate(span_t blk, buf_t* key, int (*fn)(void*, const buf_t*, const uint8_t*, size_t),
                         void* user) {
  if (blk.n < 4) return LUDA_CORRUPTION;
  uint32_t nr = dec32(blk.p + blk.n - 4);
  if (nr > (blk.n - 4) / 4) return LUDA_CORRUPTION;
  uint64_t restart_off = blk.n - (1 + (uint64_t)nr) * 4;
  if (nr == 0) return LUDA_OK; /* NewEmptyIterator */
  /* SeekToFirst (table/block.cc:202-205) starts at GetRestartPoint(0), the
   * fork's first restart word shifted right by 8 (:98-102): 0 unless the
   * first group's count spilled into the offset bits (256 or more entries),
   * in which case decoding starts mid-entry, as the fork's reader does. */
  const uint8_t* p = blk.p + (dec32(blk.p + restart_off) >> 8);
  const uint8_t* limit = blk.p + restart_off;
  key->n = 0;
  while (p < limit) {
    uint32_t shared, non_shared, vlen;
    if (limit - p < 3) return LUDA_CORRUPTION;
    shared = p[0];
    non_shared = p[1];
    vlen = p[2];
    if ((shared | non_shared | vlen) < 128) {
      p += 3;
    } else {
      if (!(p = get_varint32(p, limit, &shared))) return LUDA_CORRUPTION;
      if (!(p = get_varint32(p, limit, &non_shared))) return LUDA_CORRUPTION;
      if (!(p = get_varint32(p, limit, &vlen))) return LUDA_CORRUPTION;
    }
    if ((uint64_t)(limit - p) < (uint64_t)non_shared + vlen) return LUDA_CORRUPTION;
    if (key->n < shared) return LUDA_CORRUPTION;
    key->n = shared;
    buf_append(key, p, non_shared);
    int rc = fn(user, key, p + non_shared, vlen);
    if (rc) return rc;
    p += non_shared + vlen;
  }
  return LUDA_OK;
}

typedef struct {
  const uint8_t* data;
  uint64_t size;
  int verify;
  int (*fn)(void*, const buf_t*, const uint8_t*, size_t);
  void* user;
  buf_t dkey;
} table_ctx;

static int index_entry_cb(void* u, const buf_t* key, const uint8_t* val, size_t vlen) {
  (void)key;
  table_ctx* t = (table_ctx*)u;
  uint64_t off, n;
  const uint8_t* p = val;
  const uint8_t* lim = val + vlen;
  if (!(p = get_varint64(p, lim, &off)) || !(p = get_varint64(p, lim, &n))) return LUDA_CORRUPTION;
  span_t b;
  int rc = read_block(t->data, t->size, off, n, t->verify, &b);
  if (rc) return rc;
  return block_iterate(b, &t->dkey, t->fn, t->user);
}

static int table_walk(const uint8_t* data, uint64_t size, int verify,
                      int (*fn)(void*, const buf_t*, const uint8_t*, size_t), void* user) {
  if (size < 48) return LUDA_CORRUPTION;
  const uint8_t* f = data + size - 48;
  if (dec32(f + 40) != 0x8b80fb57u || dec32(f + 44) != 0xdb477524u) return LUDA_CORRUPTION;
  uint64_t moff, msize, ioff, isize;
  const uint8_t* p = f;
  const uint8_t* lim = f + 40;
  if (!(p = get_varint64(p, lim, &moff)) || !(p = get_varint64(p, lim, &msize)) ||
      !(p = get_varint64(p, lim, &ioff)) || !(p = get_varint64(p, lim, &isize)))
    return LUDA_CORRUPTION;
  span_t ib;
  int rc = read_block(data, size, ioff, isize, verify, &ib);
  if (rc) return rc;
  if (verify) {
    span_t mb;
    rc = read_block(data, size, moff, msize, verify, &mb);
    if (rc) return rc;
  }
  table_ctx t = {data, size, verify, fn, user, {0}};
  buf_t ikey = {0};
  rc = block_iterate(ib, &ikey, index_entry_cb, &t);
  buf_free(&ikey);
  buf_free(&t.dkey);
  return rc;
}

typedef struct {
  or_entry_fn fn;
  void* user;
  uint64_t n;
} iter_user;

static int iter_cb(void* u, const buf_t* key, const uint8_t* val, size_t vlen) {
  iter_user* it = (iter_user*)u;
  if (it->fn) it->fn(it->user, key->p, key->n, val, vlen);
  it->n++;
  return 0;
}
int or_table_iterate(const uint8_t* data, uint64_t size, int verify_crc, or_entry_fn fn, void* user,
                     uint64_t* nentries) {
  iter_user it = {fn, user, 0};
  int rc = table_walk(data, size, verify_crc, iter_cb, &it);
  if (nentries) *nentries = it.n;
  return rc;
}

/* ------------------------------------------------------------------------ */
/* build tables: db/builder.cc:17-77 with the DoCompactionWork size cut       */

int or_build_tables(const luda_options* opts, uint64_t n, const uint8_t* keys, const uint64_t* key_off,
                    const uint8_t* vals, const uint64_t* val_off, luda_result** out) {
  *out = NULL;
  if (opts->compression != 0) return LUDA_NOT_SUPPORTED;
  results_t r = {0};
  tb_t tb;
  int open = 0;
  buf_t smallest = {0}, largest = {0};
  for (uint64_t i = 0; i < n; i++) {
    const uint8_t* k = keys + key_off[i];
    size_t kl = key_off[i + 1] - key_off[i];
    if (kl < 8) {
      if (open) tb_free(&tb);
      buf_free(&r.data);
      buf_free(&r.keys);
      free(r.outs);
      buf_free(&smallest);
      buf_free(&largest);
      return LUDA_NOT_SUPPORTED;
    }
    if (!open) {
      tb_init(&tb, opts, &r.data);
      open = 1;
    }
    if (tb.num_entries == 0) buf_assign(&smallest, k, kl);
    buf_assign(&largest, k, kl);
    tb_add(&tb, k, kl, vals + val_off[i], val_off[i + 1] - val_off[i]);
    if (tb.offset >= opts->max_file_size) {
      tb_finish(&tb);
      res_add(&r, tb.base, tb.offset, (uint64_t)tb.num_entries, &smallest, &largest);
      tb_free(&tb);
      open = 0;
    }
  }
  if (open) {
    tb_finish(&tb);
    res_add(&r, tb.base, tb.offset, (uint64_t)tb.num_entries, &smallest, &largest);
    tb_free(&tb);
  }
  buf_free(&smallest);
  buf_free(&largest);
  *out = res_finish(&r, 0);
  return LUDA_OK;
}

/* ------------------------------------------------------------------------ */
/* compaction: DBImpl::DoCompactionWork (CPU), db/db_impl.cc:1060-1224        */

typedef struct {
  uint64_t* koff; /* into key arena */
  uint32_t* klen;
  const uint8_t** val;
  uint32_t* vlen;
  size_t n, cap;
  buf_t arena;
} stream_t;

static int stream_cb(void* u, const buf_t* key, const uint8_t* val, size_t vlen) {
  stream_t* s = (stream_t*)u;
  if (key->n < 8) return LUDA_NOT_SUPPORTED;
  if (s->n == s->cap) {
    s->cap = s->cap ? s->cap * 2 : 1024;
    s->koff = (uint64_t*)realloc(s->koff, s->cap * sizeof(uint64_t));
    s->klen = (uint32_t*)realloc(s->klen, s->cap * sizeof(uint32_t));
    s->val = (const uint8_t**)realloc(s->val, s->cap * sizeof(uint8_t*));
    s->vlen = (uint32_t*)realloc(s->vlen, s->cap * sizeof(uint32_t));
  }
  s->koff[s->n] = s->arena.n;
  s->klen[s->n] = (uint32_t)key->n;
  buf_append(&s->arena, key->p, key->n);
  s->val[s->n] = val;
  s->vlen[s->n] = (uint32_t)vlen;
  s->n++;
  return 0;
}
static void stream_free(stream_t* s) {
  free(s->koff);
  free(s->klen);
  free(s->val);
  free(s->vlen);
  buf_free(&s->arena);
}

/* Compaction state used by ShouldStopBefore / IsBaseLevelForKey,
 * db/version_set.h:321-391 */
typedef struct {
  const luda_job* job;
  size_t grandparent_index;
  int seen_key;
  int64_t overlapped_bytes;
  size_t level_ptrs[7];
  /* Version::files_[lvl]: the job's deeper files of that level, in order */
  size_t* level_files[7];
  size_t level_count[7];
} compaction_t;

/* Compaction::ShouldStopBefore, db/version_set.cc:1556-1579 */
static int should_stop_before(compaction_t* c, const uint8_t* key, size_t klen) {
  const luda_job* j = c->job;
  while (c->grandparent_index < j->num_grandparents &&
         or_ikey_compare(key, klen, j->grandparents[c->grandparent_index].largest.data,
                         j->grandparents[c->grandparent_index].largest.size) > 0) {
    if (c->seen_key) c->overlapped_bytes += (int64_t)j->grandparents[c->grandparent_index].file_size;
    c->grandparent_index++;
  }
  c->seen_key = 1;
  if (c->overlapped_bytes > (int64_t)(10 * j->options.max_file_size)) {
    c->overlapped_bytes = 0;
    return 1;
  }
  return 0;
}
/* Compaction::IsBaseLevelForKey, db/version_set.cc:1535-1554 */
static int is_base_level_for_key(compaction_t* c, const uint8_t* ukey, size_t un) {
  const luda_job* j = c->job;
  for (int lvl = j->level + 2; lvl < 7; lvl++) {
    while (c->level_ptrs[lvl] < c->level_count[lvl]) {
      const luda_level_file* f = &j->deeper[c->level_files[lvl][c->level_ptrs[lvl]]];
      if (bytewise_compare(ukey, un, f->largest_user_key.data, f->largest_user_key.size) <= 0) {
        if (bytewise_compare(ukey, un, f->smallest_user_key.data, f->smallest_user_key.size) >= 0)
          return 0;
        break;
      }
      c->level_ptrs[lvl]++;
    }
  }
  return 1;
}

int or_compact(const luda_job* job, luda_result** out) {
  *out = NULL;
  if (job->options.compression != 0) return LUDA_NOT_SUPPORTED;
  if (job->level < 0 || job->level >= 6) return LUDA_INVALID_ARGUMENT;

  /* VersionSet::MakeInputIterator, db/version_set.cc:1237-1268: level-0
   * inputs are one child each; every other level is one concatenated child. */
  size_t max_children = job->num_inputs[0] + 2;
  stream_t* ch = (stream_t*)calloc(max_children, sizeof(stream_t));
  size_t nch = 0;
  uint64_t bytes_read = 0;
  int rc = LUDA_OK;
  for (int which = 0; which < 2 && rc == LUDA_OK; which++) {
    if (job->num_inputs[which] == 0) continue;
    if (job->level + which == 0) {
      for (uint32_t i = 0; i < job->num_inputs[which] && rc == LUDA_OK; i++) {
        const luda_file* f = &job->inputs[which][i];
        rc = table_walk(f->data, f->size, 0, stream_cb, &ch[nch]);
        nch++;
      }
    } else {
      for (uint32_t i = 0; i < job->num_inputs[which] && rc == LUDA_OK; i++) {
        const luda_file* f = &job->inputs[which][i];
        rc = table_walk(f->data, f->size, 0, stream_cb, &ch[nch]);
      }
      nch++;
    }
    for (uint32_t i = 0; i < job->num_inputs[which]; i++) bytes_read += job->inputs[which][i].size;
  }
  if (rc != LUDA_OK) {
    for (size_t i = 0; i < max_children; i++) stream_free(&ch[i]);
    free(ch);
    return rc;
  }

  compaction_t c;
  memset(&c, 0, sizeof(c));
  c.job = job;
  for (int lvl = 0; lvl < 7; lvl++) {
    c.level_files[lvl] = (size_t*)malloc(sizeof(size_t) * (job->num_deeper ? job->num_deeper : 1));
    c.level_count[lvl] = 0;
  }
  for (uint32_t i = 0; i < job->num_deeper; i++) {
    int lvl = job->deeper[i].level;
    if (lvl >= 0 && lvl < 7) c.level_files[lvl][c.level_count[lvl]++] = i;
  }

  results_t r = {0};
  tb_t tb;
  int open = 0;
  buf_t smallest = {0}, largest = {0}, current_user_key = {0};
  int has_current_user_key = 0;
  uint64_t last_sequence_for_key = ((uint64_t)1 << 56) - 1;
  const uint64_t S = job->smallest_snapshot;
  const uint64_t max_out = job->options.max_file_size;
  size_t* pos = (size_t*)calloc(nch ? nch : 1, sizeof(size_t));
  uint64_t kept = 0, total = 0;

  for (;;) {
    /* MergingIterator::FindSmallest, table/merger.cc:148-161 */
    int best = -1;
    for (size_t i = 0; i < nch; i++) {
      if (pos[i] >= ch[i].n) continue;
      if (best < 0) {
        best = (int)i;
      } else {
        stream_t* a = &ch[i];
        stream_t* b = &ch[best];
        if (or_ikey_compare(a->arena.p + a->koff[pos[i]], a->klen[pos[i]], b->arena.p + b->koff[pos[best]],
                            b->klen[pos[best]]) < 0)
          best = (int)i;
      }
    }
    if (best < 0) break;
    stream_t* s = &ch[best];
    size_t idx = pos[best]++;
    const uint8_t* key = s->arena.p + s->koff[idx];
    size_t klen = s->klen[idx];
    total++;

    if (should_stop_before(&c, key, klen) && open) {
      tb_finish(&tb);
      res_add(&r, tb.base, tb.offset, (uint64_t)tb.num_entries, &smallest, &largest);
      tb_free(&tb);
      open = 0;
    }

    int drop = 0;
    uint64_t tag = dec64(key + klen - 8);
    uint8_t type = (uint8_t)(tag & 0xff);
    uint64_t seq = tag >> 8;
    if (type > 1) { /* ParseInternalKey fails, db/dbformat.h:165-175 */
      current_user_key.n = 0;
      has_current_user_key = 0;
      last_sequence_for_key = ((uint64_t)1 << 56) - 1;
    } else {
      if (!has_current_user_key ||
          bytewise_compare(key, klen - 8, current_user_key.p, current_user_key.n) != 0) {
        buf_assign(&current_user_key, key, klen - 8);
        has_current_user_key = 1;
        last_sequence_for_key = ((uint64_t)1 << 56) - 1;
      }
      if (last_sequence_for_key <= S) {
        drop = 1;
      } else if (type == 0 && seq <= S && is_base_level_for_key(&c, key, klen - 8)) {
        drop = 1;
      }
      last_sequence_for_key = seq;
    }

    if (!drop) {
      if (!open) {
        tb_init(&tb, &job->options, &r.data);
        open = 1;
      }
      if (tb.num_entries == 0) buf_assign(&smallest, key, klen);
      buf_assign(&largest, key, klen);
      tb_add(&tb, key, klen, s->val[idx], s->vlen[idx]);
      kept++;
      if (tb.offset >= max_out) {
        tb_finish(&tb);
        res_add(&r, tb.base, tb.offset, (uint64_t)tb.num_entries, &smallest, &largest);
        tb_free(&tb);
        open = 0;
      }
    }
  }
  if (open) {
    tb_finish(&tb);
    res_add(&r, tb.base, tb.offset, (uint64_t)tb.num_entries, &smallest, &largest);
    tb_free(&tb);
  }
  buf_free(&smallest);
  buf_free(&largest);
  buf_free(&current_user_key);
  free(pos);
  for (int lvl = 0; lvl < 7; lvl++) free(c.level_files[lvl]);
  for (size_t i = 0; i < max_children; i++) stream_free(&ch[i]);
  free(ch);
  *out = res_finish(&r, bytes_read);
  (*out)->stats.entries_in = total;
  (*out)->stats.entries_kept = kept;
  return LUDA_OK;
}

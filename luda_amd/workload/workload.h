/*
 * workload.h — seeded synthetic compaction jobs (host only, no GPU code).
 *
 * A workload is a compaction job described as sorted streams of internal
 * keys and values ("groups").  Each group becomes one or more input tables:
 * the caller encodes it with a table builder (the engine's
 * luda_build_tables, or a CPU builder in tests), cutting files at
 * max_file_size exactly as DoCompactionWork does, and lists the tables as
 * Compaction::inputs_[which].  Grandparent and deeper-level metadata go into
 * luda_job unchanged.
 *
 * Job shapes:
 *   0  L0->L1, 4 overlapping L0 tables + 6 L1 tables, 16-byte decimal keys
 *      mod 1e6, 100-byte values, 20% of L0 keys overwrite L1 keys, 2% deletes
 *   1  L1->L2, one L1 table + ~115 L2 tables (256 MiB at scale 1), keys uniform
 *      over 1e8, 10% overwrites, 3% tombstones, not bottommost (an L3 file
 *      covers the key space, so tombstones stay and no grandparent cut fires)
 *   2  bottommost L1->L2, 512 MiB, Zipf(0.99) ids over 1e7, 8-64 byte keys with
 *      an 8-byte prefix per 4096-id range, 16-1024 byte values, 5% deletes,
 *      snapshots at the 25/50/75% sequence quantiles
 *   3  L0->L1 with 4-12 level-0 tables (the seed picks how many), the
 *      distribution of job 1 (256 MiB at scale 1)
 *   4  the same with the distribution of job 2
 *   5  grandparent-cut stress: L1->L2 with many small L3 grandparents so
 *      ShouldStopBefore fires, plus deeper-level ranges that make
 *      IsBaseLevelForKey vary
 * `scale` multiplies every entry count (1.0 = the full size).
 */
#ifndef LUDA_WORKLOAD_H_
#define LUDA_WORKLOAD_H_

#include <stdint.h>

#include "../../include/luda.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lw_group {
  int32_t which;          /* Compaction::inputs_[which] */
  uint32_t reserved;
  uint64_t max_file_size; /* cut into tables at this size; UINT64_MAX = one table */
  uint64_t n;
  uint8_t* keys;          /* internal keys, entry i = keys[key_off[i] .. key_off[i+1]) */
  uint64_t* key_off;
  uint8_t* vals;
  uint64_t* val_off;
} lw_group;

typedef struct lw_workload {
  int32_t level;
  uint32_t num_groups;
  uint64_t smallest_snapshot;
  lw_group* groups;
  uint32_t num_grandparents;
  uint32_t num_deeper;
  luda_grandparent* grandparents;
  luda_level_file* deeper;
  uint8_t* meta_keys; /* backing store of grandparent / deeper keys */
  uint64_t total_entries;
} lw_workload;

lw_workload* lw_make(int config, uint64_t seed, double scale);
void lw_free(lw_workload* w);

#ifdef __cplusplus
}
#endif

#endif /* LUDA_WORKLOAD_H_ */

// workload.cc — seeded synthetic compaction jobs (see workload.h).
//
// Keys follow db/db_bench.cc:729 ("%016llu" zero-padded decimal); internal
// keys append fixed64(seq << 8 | type) (db/dbformat.cc:20-23).  The PRNG is
// splitmix64 (64-bit, unlike LevelDB's 31-bit util/random.h).
#include "workload.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <unordered_set>
#include <vector>

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t Next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double Uniform() { return (Next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t Below(uint64_t n) { return n ? Next() % n : 0; }
};

struct Entry {
  std::string ukey;
  uint64_t seq;
  uint8_t type;  // 1 value, 0 deletion
  uint32_t vlen;
  uint64_t vseed;
};

// Internal-key order: user key ascending, then sequence descending
// (db/dbformat.cc:51-67; the type breaks no ties here because every entry
// has its own sequence number).
bool EntryLess(const Entry& a, const Entry& b) {
  int r = a.ukey.compare(b.ukey);
  if (r != 0) return r < 0;
  return a.seq > b.seq;
}

std::string Decimal16(uint64_t v) {
  char buf[32];
  snprintf(buf, sizeof(buf), "%016llu", (unsigned long long)v);
  return std::string(buf, 16);
}

void PutFixed64(std::string* s, uint64_t v) {
  for (int i = 0; i < 8; i++) s->push_back((char)(v >> (8 * i)));
}

struct Builder {
  lw_workload* w;
  std::vector<lw_group> groups;
  std::string meta;
  std::vector<luda_grandparent> gps;
  std::vector<std::pair<size_t, size_t>> gp_keys;  // (offset, size) into meta
  struct Deep {
    size_t so, sn, lo, ln;
    int level;
  };
  std::vector<Deep> deep;

  void AddGroup(int which, uint64_t max_file_size, std::vector<Entry>& es) {
    std::sort(es.begin(), es.end(), EntryLess);
    lw_group g;
    memset(&g, 0, sizeof(g));
    g.which = which;
    g.max_file_size = max_file_size;
    g.n = es.size();
    uint64_t kbytes = 0, vbytes = 0;
    for (auto& e : es) {
      kbytes += e.ukey.size() + 8;
      vbytes += e.vlen;
    }
    g.keys = (uint8_t*)malloc(kbytes ? kbytes : 1);
    g.vals = (uint8_t*)malloc(vbytes ? vbytes : 1);
    g.key_off = (uint64_t*)malloc(sizeof(uint64_t) * (g.n + 1));
    g.val_off = (uint64_t*)malloc(sizeof(uint64_t) * (g.n + 1));
    uint64_t ko = 0, vo = 0;
    for (uint64_t i = 0; i < g.n; i++) {
      const Entry& e = es[i];
      g.key_off[i] = ko;
      memcpy(g.keys + ko, e.ukey.data(), e.ukey.size());
      ko += e.ukey.size();
      uint64_t tag = (e.seq << 8) | e.type;
      for (int b = 0; b < 8; b++) g.keys[ko++] = (uint8_t)(tag >> (8 * b));
      g.val_off[i] = vo;
      Rng r(e.vseed);
      uint32_t j = 0;
      for (; j + 8 <= e.vlen; j += 8) {
        uint64_t x = r.Next();
        memcpy(g.vals + vo + j, &x, 8);
      }
      if (j < e.vlen) {
        uint64_t x = r.Next();
        memcpy(g.vals + vo + j, &x, e.vlen - j);
      }
      vo += e.vlen;
    }
    g.key_off[g.n] = ko;
    g.val_off[g.n] = vo;
    groups.push_back(g);
  }
  void AddGrandparent(const std::string& largest_ikey, uint64_t size) {
    gp_keys.push_back({meta.size(), largest_ikey.size()});
    meta += largest_ikey;
    luda_grandparent g;
    memset(&g, 0, sizeof(g));
    g.file_size = size;
    gps.push_back(g);
  }
  void AddDeeper(int level, const std::string& smallest, const std::string& largest) {
    Deep d;
    d.so = meta.size();
    d.sn = smallest.size();
    meta += smallest;
    d.lo = meta.size();
    d.ln = largest.size();
    meta += largest;
    d.level = level;
    deep.push_back(d);
  }
  lw_workload* Finish(int level, uint64_t snapshot) {
    lw_workload* out = (lw_workload*)calloc(1, sizeof(lw_workload));
    out->level = level;
    out->smallest_snapshot = snapshot;
    out->num_groups = (uint32_t)groups.size();
    out->groups = (lw_group*)malloc(sizeof(lw_group) * (groups.size() ? groups.size() : 1));
    for (size_t i = 0; i < groups.size(); i++) {
      out->groups[i] = groups[i];
      out->total_entries += groups[i].n;
    }
    out->meta_keys = (uint8_t*)malloc(meta.size() ? meta.size() : 1);
    memcpy(out->meta_keys, meta.data(), meta.size());
    out->num_grandparents = (uint32_t)gps.size();
    out->grandparents = (luda_grandparent*)malloc(sizeof(luda_grandparent) * (gps.size() ? gps.size() : 1));
    for (size_t i = 0; i < gps.size(); i++) {
      out->grandparents[i] = gps[i];
      out->grandparents[i].largest.data = out->meta_keys + gp_keys[i].first;
      out->grandparents[i].largest.size = gp_keys[i].second;
    }
    out->num_deeper = (uint32_t)deep.size();
    out->deeper = (luda_level_file*)calloc(deep.size() ? deep.size() : 1, sizeof(luda_level_file));
    for (size_t i = 0; i < deep.size(); i++) {
      out->deeper[i].smallest_user_key.data = out->meta_keys + deep[i].so;
      out->deeper[i].smallest_user_key.size = deep[i].sn;
      out->deeper[i].largest_user_key.data = out->meta_keys + deep[i].lo;
      out->deeper[i].largest_user_key.size = deep[i].ln;
      out->deeper[i].level = deep[i].level;
    }
    return out;
  }
};

uint64_t Scaled(double base, double scale) {
  double v = base * scale;
  return v < 1 ? 1 : (uint64_t)v;
}

// ---- job 0: L0 -> L1 -------------------------------------------------------
lw_workload* MakeL0(uint64_t seed, double scale) {
  Rng rng(seed);
  Builder b;
  const uint64_t kSpace = 1000000;
  const uint64_t n1 = Scaled(6 * 18000, scale);  // 6 tables of ~2 MiB
  const uint64_t per_l0 = Scaled(17900, scale);
  std::vector<uint64_t> l1ids;
  {
    std::unordered_set<uint64_t> seen;
    while (l1ids.size() < n1 && l1ids.size() < kSpace) {
      uint64_t id = rng.Next() % kSpace;
      if (seen.insert(id).second) l1ids.push_back(id);
    }
  }
  uint64_t seq = 0;
  std::vector<Entry> l1;
  std::vector<uint64_t> order(l1ids.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = i;
  for (size_t i = order.size(); i > 1; i--) std::swap(order[i - 1], order[rng.Below(i)]);
  l1.resize(l1ids.size());
  for (size_t i = 0; i < l1ids.size(); i++) {
    Entry& e = l1[order[i]];
    e.ukey = Decimal16(l1ids[order[i]]);
    e.seq = ++seq;
    e.type = 1;
    e.vlen = 100;
    e.vseed = rng.Next();
  }
  // L0 tables: file 0 oldest.
  std::vector<std::vector<Entry>> l0(4);
  for (int f = 0; f < 4; f++) {
    for (uint64_t i = 0; i < per_l0; i++) {
      Entry e;
      uint64_t id = (rng.Uniform() < 0.2 && !l1ids.empty()) ? l1ids[rng.Below(l1ids.size())] : rng.Next() % kSpace;
      e.ukey = Decimal16(id);
      e.seq = ++seq;
      bool del = rng.Uniform() < 0.02;
      e.type = del ? 0 : 1;
      e.vlen = del ? 0 : 100;
      e.vseed = rng.Next();
      l0[f].push_back(e);
    }
  }
  for (int f = 0; f < 4; f++) b.AddGroup(0, UINT64_MAX, l0[f]);
  b.AddGroup(1, 2 << 20, l1);
  return b.Finish(0, seq);
}

// ---- job 1: L1 -> L2, not bottommost ---------------------------------------
lw_workload* MakeL1(uint64_t seed, double scale) {
  Rng rng(seed);
  Builder b;
  const uint64_t kSpace = 100000000;
  const uint64_t n2 = Scaled(2030000, scale);
  const uint64_t n1 = Scaled(220000, scale);
  std::vector<uint64_t> ids2;
  {
    std::unordered_set<uint64_t> seen;
    seen.reserve(n2 * 2);
    while (ids2.size() < n2) {
      uint64_t id = rng.Next() % kSpace;
      if (seen.insert(id).second) ids2.push_back(id);
    }
  }
  std::vector<uint64_t> ids1;
  {
    std::unordered_set<uint64_t> seen;
    seen.reserve(n1 * 2);
    while (ids1.size() < n1) {
      uint64_t id = rng.Uniform() < 0.1 ? ids2[rng.Below(ids2.size())] : rng.Next() % kSpace;
      if (seen.insert(id).second) ids1.push_back(id);
    }
  }
  uint64_t seq = 0;
  std::vector<Entry> l2(ids2.size()), l1(ids1.size());
  for (size_t i = 0; i < ids2.size(); i++) {  // insertion order is the random draw order
    l2[i].ukey = Decimal16(ids2[i]);
    l2[i].seq = ++seq;
    l2[i].type = 1;
    l2[i].vlen = 100;
    l2[i].vseed = rng.Next();
  }
  for (size_t i = 0; i < ids1.size(); i++) {
    l1[i].ukey = Decimal16(ids1[i]);
    l1[i].seq = ++seq;
    bool del = rng.Uniform() < 0.03;
    l1[i].type = del ? 0 : 1;
    l1[i].vlen = del ? 0 : 100;
    l1[i].vseed = rng.Next();
  }
  b.AddGroup(0, UINT64_MAX, l1);
  b.AddGroup(1, 2 << 20, l2);
  // One L3 file spanning the whole key space: IsBaseLevelForKey is false for
  // every key (tombstones kept) and no key passes its largest key, so
  // ShouldStopBefore never fires.
  std::string lo = Decimal16(0), hi = Decimal16(9999999999999999ull);
  std::string hi_ikey = hi;
  PutFixed64(&hi_ikey, (1ull << 8) | 1);
  b.AddGrandparent(hi_ikey, 1ull << 30);
  b.AddDeeper(3, lo, hi);
  return b.Finish(1, seq);
}

// Zipf(theta) over [0, n) by Gray et al., "Quickly generating billion-record
// synthetic databases" (SIGMOD 1994).
struct Zipf {
  uint64_t n;
  double theta, alpha, zetan, eta;
  Zipf(uint64_t n_, double th) : n(n_), theta(th) {
    double z = 0;
    for (uint64_t i = 1; i <= n; i++) z += 1.0 / pow((double)i, theta);
    zetan = z;
    double zeta2 = 1.0 + 1.0 / pow(2.0, theta);
    alpha = 1.0 / (1.0 - theta);
    eta = (1 - pow(2.0 / n, 1 - theta)) / (1 - zeta2 / zetan);
  }
  uint64_t Next(Rng& r) const {
    double u = r.Uniform();
    double uz = u * zetan;
    if (uz < 1.0) return 0;
    if (uz < 1.0 + pow(0.5, theta)) return 1;
    uint64_t v = (uint64_t)(n * pow(eta * u - eta + 1, alpha));
    return v >= n ? n - 1 : v;
  }
};

// User key of id: 8-byte big-endian prefix floor(id/4096), then big-endian
// (id % 4096) and id-seeded filler, cut to a length in [8, 64].
std::string SkewKey(uint64_t id) {
  Rng r(id * 0x9E3779B97F4A7C15ull + 12345);
  uint32_t len = 8 + (uint32_t)(r.Next() % 57);
  std::string k;
  uint64_t pre = id / 4096;
  for (int i = 7; i >= 0; i--) k.push_back((char)(pre >> (8 * i)));
  uint32_t lo = (uint32_t)(id % 4096);
  k.push_back((char)(lo >> 8));
  k.push_back((char)lo);
  while (k.size() < len) k.push_back((char)r.Next());
  k.resize(len);
  return k;
}

// ---- job 2: skewed, multi-version, bottommost -------------------------------
lw_workload* MakeSkewed(uint64_t seed, double scale) {
  Rng rng(seed);
  Builder b;
  const uint64_t m = Scaled(950000, scale);
  const uint64_t nid = scale >= 0.999 ? 10000000ull : std::max<uint64_t>(1000, (uint64_t)(10000000ull * scale));
  Zipf z(nid, 0.99);
  std::vector<Entry> all(m);
  for (uint64_t i = 0; i < m; i++) {
    uint64_t id = z.Next(rng);
    // Scatter the hot ids over the key space.
    id = (id * 2654435761ull) % nid;
    Entry& e = all[i];
    e.ukey = SkewKey(id);
    e.seq = i + 1;
    bool del = rng.Uniform() < 0.05;
    e.type = del ? 0 : 1;
    e.vlen = del ? 0 : 16 + (uint32_t)rng.Below(1009);
    e.vseed = rng.Next();
  }
  uint64_t cut = (uint64_t)(0.9 * m);
  std::vector<Entry> older(all.begin(), all.begin() + cut), newer(all.begin() + cut, all.end());
  b.AddGroup(0, 2 << 20, newer);
  b.AddGroup(1, 2 << 20, older);
  // Snapshots at the 25/50/75% quantiles: only the oldest matters.
  uint64_t snap = std::max<uint64_t>(1, m / 4);
  return b.Finish(1, snap);
}

// ---- jobs 3 and 4: L0 -> L1 with 4-12 level-0 files --------------------------
// The mixed batch's level-0 jobs.  The seed picks the number of level-0 files
// k in [4, 12]; 40% of the entries are the newest, split by sequence into k
// memtable-sized runs (one table each, as flushes write them), the older 60%
// form the level-1 input cut at 2 MiB.  Job 3 uses the 256 MiB job's key and
// value distribution (16-byte decimal keys uniform over 1e8, 100-byte values,
// 2% deletes, no snapshot); job 4 the skewed one (Zipf(0.99) ids, 8-64 byte
// keys, 16-1024 byte values, 5% deletes, oldest snapshot at the 25% quantile).
// scale 1.0 is 256 MiB of input; no deeper level, so tombstones older than
// the snapshot are dropped.
lw_workload* MakeL0Many(uint64_t seed, double scale, bool skewed) {
  Rng rng(seed);
  Builder b;
  const uint32_t k = 4 + (uint32_t)rng.Below(9);
  const uint64_t m = Scaled(skewed ? 475000 : 2250000, scale);
  std::vector<Entry> all(m);
  const uint64_t nid = skewed ? std::max<uint64_t>(1000, (uint64_t)(10000000ull * std::min(1.0, scale))) : 0;
  Zipf* z = skewed ? new Zipf(nid, 0.99) : nullptr;
  for (uint64_t i = 0; i < m; i++) {
    Entry& e = all[i];
    bool del;
    if (skewed) {
      const uint64_t id = (z->Next(rng) * 2654435761ull) % nid;
      e.ukey = SkewKey(id);
      del = rng.Uniform() < 0.05;
      e.vlen = del ? 0 : 16 + (uint32_t)rng.Below(1009);
    } else {
      e.ukey = Decimal16(rng.Next() % 100000000ull);
      del = rng.Uniform() < 0.02;
      e.vlen = del ? 0 : 100;
    }
    e.seq = i + 1;
    e.type = del ? 0 : 1;
    e.vseed = rng.Next();
  }
  delete z;
  const uint64_t cut = (uint64_t)(0.6 * m);
  std::vector<Entry> older(all.begin(), all.begin() + cut);
  // A level holds each user key once per version, sorted; no dedupe is
  // needed for a compaction input (DoCompactionWork drops the shadowed ones).
  for (uint32_t f = 0; f < k; f++) {
    const uint64_t a = cut + (m - cut) * f / k, e = cut + (m - cut) * (f + 1) / k;
    std::vector<Entry> run(all.begin() + a, all.begin() + e);
    b.AddGroup(0, UINT64_MAX, run);
  }
  b.AddGroup(1, 2 << 20, older);
  return b.Finish(0, skewed ? std::max<uint64_t>(1, m / 4) : m);
}

// ---- job 5: grandparent cuts and a patchy deeper level ----------------------
lw_workload* MakeCuts(uint64_t seed, double scale) {
  Rng rng(seed);
  Builder b;
  const uint64_t kSpace = 1000000;
  const uint64_t n2 = Scaled(60000, scale), n1 = Scaled(20000, scale);
  std::vector<Entry> l1, l2;
  uint64_t seq = 0;
  for (uint64_t i = 0; i < n2; i++) {
    Entry e;
    e.ukey = Decimal16(rng.Next() % kSpace);
    e.seq = ++seq;
    e.type = 1;
    e.vlen = 20 + (uint32_t)rng.Below(200);
    e.vseed = rng.Next();
    l2.push_back(e);
  }
  for (uint64_t i = 0; i < n1; i++) {
    Entry e;
    e.ukey = Decimal16(rng.Next() % kSpace);
    e.seq = ++seq;
    bool del = rng.Uniform() < 0.3;
    e.type = del ? 0 : 1;
    e.vlen = del ? 0 : 20 + (uint32_t)rng.Below(200);
    e.vseed = rng.Next();
    l1.push_back(e);
  }
  // Dedupe each level to one version per user key (newest wins).
  auto dedupe = [](std::vector<Entry>& v) {
    std::sort(v.begin(), v.end(), EntryLess);
    std::vector<Entry> out;
    for (auto& e : v)
      if (out.empty() || out.back().ukey != e.ukey) out.push_back(e);
    v.swap(out);
  };
  dedupe(l1);
  dedupe(l2);
  b.AddGroup(0, 1 << 20, l1);
  b.AddGroup(1, 1 << 20, l2);
  // 200 grandparent files of 1.5 MiB each partitioning the key space: the
  // overlap budget (10 x max_file_size) is crossed every ~7 files.
  for (int i = 1; i <= 200; i++) {
    std::string k = Decimal16((uint64_t)i * kSpace / 200 - 1);
    PutFixed64(&k, (5ull << 8) | 1);
    b.AddGrandparent(k, 3 << 19);
  }
  // Deeper levels: L3 covers every other 1/50 slice of the key space; L4 one
  // range; so IsBaseLevelForKey is true for some tombstones and false for others.
  for (int i = 0; i < 50; i += 2)
    b.AddDeeper(3, Decimal16((uint64_t)i * kSpace / 50), Decimal16((uint64_t)(i + 1) * kSpace / 50 - 1));
  b.AddDeeper(4, Decimal16(kSpace / 3), Decimal16(kSpace / 3 + 5000));
  return b.Finish(1, seq - n1 / 2);
}

}  // namespace

extern "C" {

lw_workload* lw_make(int config, uint64_t seed, double scale) {
  switch (config) {
    case 0:
      return MakeL0(seed, scale);
    case 1:
      return MakeL1(seed, scale);
    case 2:
      return MakeSkewed(seed, scale);
    case 3:
      return MakeL0Many(seed, scale, false);
    case 4:
      return MakeL0Many(seed, scale, true);
    case 5:
      return MakeCuts(seed, scale);
    default:
      return nullptr;
  }
}

void lw_free(lw_workload* w) {
  if (!w) return;
  for (uint32_t i = 0; i < w->num_groups; i++) {
    free(w->groups[i].keys);
    free(w->groups[i].vals);
    free(w->groups[i].key_off);
    free(w->groups[i].val_off);
  }
  free(w->groups);
  free(w->grandparents);
  free(w->deeper);
  free(w->meta_keys);
  free(w);
}

}  // extern "C"

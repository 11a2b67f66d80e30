// Host build of the block writer's value staging (luda_amd/csrc/stage_values.h,
// used by k_write_blocks in k_encode.hip): random data blocks are laid out in
// a stage the way the kernel does it (per round of up to 64 entries: the
// values first, as aligned dwords with junk at their ends or by 16-byte
// slots; then the headers and keys byte-exact; then the restart words; after
// the last round the first restart word again, the count and the type byte),
// and the stage is compared with the block as BlockBuilder lays it out
// (table/block_builder.cc:62-114: entries, restart array, count) plus the
// type byte.  Every stage write and every value read is bounds-checked: a
// value's source allows reads up to 15 bytes past its end.
//
// usage: stage_harness SEED BLOCKS   prints "ok bytes=N" (exit 0) or the first mismatch (exit 1)
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "stage_values.h"

using std::vector;

static constexpr uint32_t kStage = 6144;  // k_encode.hip: bytes of one block's stage per wave
static int fail(const char* what, long a, long b) {
  printf("FAIL %s (%ld, %ld)\n", what, a, b);
  exit(1);
}

struct Entry {
  vector<uint8_t> head;  // header varints and key delta
  vector<uint8_t> val;   // the value, followed by 15 readable bytes
  uint32_t vl;
};

static void put_varint(vector<uint8_t>& o, uint32_t v) {
  while (v >= 128) {
    o.push_back((uint8_t)(v | 128));
    v >>= 7;
  }
  o.push_back((uint8_t)v);
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  std::mt19937_64 rng(strtoull(argv[1], nullptr, 10));
  const int blocks = atoi(argv[2]);
  uint64_t total = 0;
  for (int bi = 0; bi < blocks; bi++) {
    auto rnd = [&](uint64_t n) { return n ? rng() % n : 0; };
    const uint32_t R = 1 + (uint32_t)rnd(rnd(2) ? 16 : 40);
    const int shape = (int)rnd(4);  // 0 small values, 1 large values, 2 mixed, 3 tiny
    const uint32_t want = 1 + (uint32_t)rnd(shape == 3 ? 200 : 90);
    vector<Entry> es;
    uint64_t bytes = 0;
    for (uint32_t j = 0; j < want; j++) {
      Entry e;
      uint32_t vl;
      if (shape == 0) vl = (uint32_t)rnd(129);
      else if (shape == 1) vl = 129 + (uint32_t)rnd(1000);
      else if (shape == 2) vl = rnd(3) ? (uint32_t)rnd(129) : 129 + (uint32_t)rnd(600);
      else vl = (uint32_t)rnd(6);
      const uint32_t sh = (j % R == 0) ? 0 : (uint32_t)rnd(rnd(4) ? 20 : 300);
      const uint32_t ns = (uint32_t)rnd(shape == 3 ? 3 : 30);
      put_varint(e.head, sh);
      put_varint(e.head, ns);
      put_varint(e.head, vl);
      for (uint32_t t = 0; t < ns; t++) e.head.push_back((uint8_t)rng());
      e.vl = vl;
      e.val.resize(vl + 15);
      for (auto& x : e.val) x = (uint8_t)rng();
      const uint64_t add = e.head.size() + vl;
      if (bytes + add + 4ull * ((j + R) / R) + 4 + 5 + 16 + 15 > kStage) break;
      bytes += add;
      es.push_back(std::move(e));
    }
    if (es.empty()) continue;
    const uint32_t cnt = (uint32_t)es.size();
    const uint32_t ngroups = (cnt + R - 1) / R;
    const uint64_t dend = bytes;
    const uint64_t bsz = dend + 4ull * ngroups + 4;
    // BlockBuilder's layout
    vector<uint8_t> ref;
    vector<uint32_t> off(cnt), voff(cnt);
    for (uint32_t j = 0; j < cnt; j++) {
      off[j] = (uint32_t)ref.size();
      ref.insert(ref.end(), es[j].head.begin(), es[j].head.end());
      voff[j] = (uint32_t)ref.size();
      ref.insert(ref.end(), es[j].val.begin(), es[j].val.begin() + es[j].vl);
    }
    for (uint32_t g = 0; g < ngroups; g++) {
      const uint32_t j = g * R, w = (off[j] << 8) | (R < cnt - j ? R : cnt - j);
      for (int t = 0; t < 4; t++) ref.push_back((uint8_t)(w >> (8 * t)));
    }
    for (int t = 0; t < 4; t++) ref.push_back((uint8_t)(ngroups >> (8 * t)));
    ref.push_back(0);
    if (ref.size() != bsz + 1) fail("reference size", (long)ref.size(), (long)bsz);

    // the stage: kStage + 32 bytes, 16-byte aligned, the block at mis
    const uint32_t mis = (uint32_t)rnd(16);
    vector<uint8_t> stage(kStage + 32);
    for (auto& x : stage) x = (uint8_t)rng();
    auto w8 = [&](uint64_t x, uint8_t v) {  // x: stage offset of the block's byte
      if (mis + x >= stage.size()) fail("stage write", (long)(mis + x), (long)stage.size());
      stage[mis + x] = v;
    };
    const bool lane_path_kh2 = rnd(2);
    for (uint32_t j0 = 0; j0 < cnt; j0 += 64) {
      const uint32_t j1 = j0 + 64 < cnt ? j0 + 64 : cnt;
      uint32_t maxch = 0;
      for (uint32_t j = j0; j < j1; j++) maxch = std::max(maxch, (es[j].vl + 15) >> 4);
      if (maxch <= 8) {
        for (uint32_t j = j0; j < j1; j++) {
          const Entry& e = es[j];
          auto chunk = [&](uint32_t t) {
            if (16ull * t + 16 > e.vl + 15ull) fail("value read", (long)t, (long)e.vl);
            sv::Chunk c;
            memcpy(c.w, e.val.data() + 16 * t, 16);
            return c;
          };
          const uint32_t at = mis + voff[j];  // stage byte address (the stage is 16-byte aligned)
          auto put = [&](uint32_t k, uint32_t w) {
            const uint64_t a = (at & ~3u) + 4ull * k;
            for (int t = 0; t < 4; t++) w8(a - mis + t, (uint8_t)(w >> (8 * t)));
          };
          const uint32_t nch = (e.vl + 15) >> 4;
          if (lane_path_kh2) {
            sv::Chunk first[2];
            for (uint32_t t = 0; t < 2; t++)
              if (t < nch) first[t] = chunk(t);
            sv::store_value<2>(at, e.vl, first, chunk, put);
          } else {
            sv::Chunk first[4];
            for (uint32_t t = 0; t < 4; t++)
              if (t < nch) first[t] = chunk(t);
            sv::store_value<4>(at, e.vl, first, chunk, put);
          }
        }
      } else {
        for (uint32_t j = j0; j < j1; j++) {
          const Entry& e = es[j];
          if (!e.vl) continue;
          const uint32_t ns = sv::spread_slots(mis, voff[j], e.vl);
          for (uint32_t k = 0; k < ns; k++) {
            const sv::Slot s = sv::spread_slot(mis, (int)voff[j], (int)e.vl, k);
            if (s.hi <= s.lo) fail("empty slot", s.lo, s.hi);
            const long r = s.lo - (long)voff[j];
            if (r < 0 || r + 16 > (long)e.vl + 15) fail("spread read", r, e.vl);
            if (s.full && ((mis + (uint32_t)s.x) & 15)) fail("unaligned slot", s.x, mis);
            for (int t = s.lo; t < s.hi; t++) w8((uint64_t)t, e.val[r + (t - s.lo)]);
          }
        }
      }
      // headers and keys, then this round's restart words
      for (uint32_t j = j0; j < j1; j++)
        for (uint32_t t = 0; t < es[j].head.size(); t++) w8(off[j] + t, es[j].head[t]);
      for (uint32_t j = j0; j < j1; j++)
        if (j % R == 0) {
          const uint32_t w = (off[j] << 8) | (R < cnt - j ? R : cnt - j);
          for (int t = 0; t < 4; t++) w8(dend + 4ull * (j / R) + t, (uint8_t)(w >> (8 * t)));
        }
    }
    {
      const uint32_t w0 = R < cnt ? R : cnt;  // restart word 0 again (entry 0 at offset 0)
      for (int t = 0; t < 4; t++) w8(dend + t, (uint8_t)(w0 >> (8 * t)));
      for (int t = 0; t < 4; t++) w8(dend + 4ull * ngroups + t, (uint8_t)(ngroups >> (8 * t)));
      w8(bsz, 0);
    }
    for (uint64_t x = 0; x <= bsz; x++)
      if (stage[mis + x] != ref[x]) {
        printf("FAIL block %d byte %lu of %lu: stage %02x want %02x (mis %u, R %u, entries %u)\n", bi,
               (unsigned long)x, (unsigned long)bsz, stage[mis + x], ref[x], mis, R, cnt);
        return 1;
      }
    total += bsz + 1;
  }
  printf("ok bytes=%lu\n", (unsigned long)total);
  return 0;
}

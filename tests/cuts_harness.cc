// Host build of the tiled block cuts (luda_amd/csrc/cuts.h) checked against a
// one-entry-at-a-time restatement of TableBuilder's block flush
// (table/table_builder.cc:119-123; block_builder.cc:62-114 for the estimate)
// with DoCompactionWork's cut before a ShouldStopBefore entry
// (db/db_impl.cc:1105-1111).  The tiles, their entry-point functions, the
// composition from the first tile on and the emitted blocks follow the
// device kernels (k_cuts.hip) step by step, with every array index checked.
//
// usage: cuts_harness SEED M R BS SHAPE FORCED_EVERY
//   SHAPE: fixed | varied | small | huge | mixed
// prints "ok blocks=N fallback=F" (exit 0) or the first mismatch (exit 1).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "cuts.h"

using std::vector;

struct Block {
  uint32_t first, n, bytes, forced;
  bool operator!=(const Block& o) const {
    return first != o.first || n != o.n || bytes != o.bytes || forced != o.forced;
  }
};

// TableBuilder::Add then "estimate >= block_size: Flush", a table (so a
// block) ending before a forced entry, and the last block at the stream end.
static vector<Block> reference(const vector<uint32_t>& szr, const vector<uint32_t>& sznr,
                               const vector<uint32_t>& forced, uint64_t bs, uint32_t R) {
  vector<Block> out;
  const uint32_t m = (uint32_t)szr.size();
  uint32_t x = 0;
  while (x < m) {
    uint64_t bytes = 0, g = 0, est = 0;
    uint32_t y = x;
    for (;;) {
      const bool rs = (y - x) % R == 0;
      bytes += rs ? szr[y] : sznr[y];
      g += rs;
      y++;
      est = bytes + 4 * g + 4;
      if (est >= bs || y == m || forced[y]) break;
    }
    out.push_back(Block{x, y - x, (uint32_t)est, (x > 0 && forced[x]) ? 1u : 0u});
    x = y;
  }
  return out;
}

struct TileOut {
  uint32_t te, lim, C;
  vector<uint32_t> ftab;
};

static cuts::Tile g_tile;
static vector<uint32_t> g_bb;  // block bytes of each start (bbytes of cut_tiles)

// cut_tiles of k_cuts.hip: the staged entries, their prefixes, every start's
// block end and bytes.  Returns false when the tile is left to the orbit
// kernels (a block end unresolved, bytes or surcharges too wide).
static bool stage(const vector<uint32_t>& szr, const vector<uint32_t>& sznr, const vector<uint32_t>& forced,
                  uint64_t T0, uint64_t bs, uint32_t R, uint32_t* te, uint32_t* lim) {
  const uint64_t m = szr.size();
  cuts::Tile& L = g_tile;
  *te = (uint32_t)std::min<uint64_t>(cuts::kTile, m - T0);
  *lim = (uint32_t)std::min<uint64_t>(cuts::kStaged, m - T0);
  bool ok = true;
  uint64_t run = 0;
  uint32_t frun = 0;
  L.P[0] = 0;
  L.F[0] = 0;
  for (uint32_t x = 0; x < *lim; x++) {
    const uint64_t q = T0 + x;
    if (szr[q] - sznr[q] > 0xffffu) ok = false;
    L.D[x] = (uint16_t)(szr[q] - sznr[q]);
    run += sznr[q];
    frun += forced[q] ? 1 : 0;
    if (run > 0xffffffffull) ok = false;
    L.P[x + 1] = (uint32_t)run;
    L.F[x + 1] = (uint16_t)frun;
  }
  const bool stream_end = T0 + *lim == m;
  const uint32_t est = cuts::mean_block(L, *lim, bs);
  g_bb.assign(cuts::kTile, 0);
  for (uint32_t x = 0; x < *te; x++) {
    uint64_t bytes = 0;
    uint32_t e = cuts::block_end(L, x, *lim, stream_end, est, bs, R, &bytes);
    if (e == cuts::kUnresolved) {
      ok = false;
      e = x + 1;
    }
    L.nx[x] = (uint16_t)e;
    g_bb[x] = (uint32_t)bytes;
  }
  return ok;
}

int main(int argc, char** argv) {
  if (argc != 7) {
    fprintf(stderr, "usage: cuts_harness SEED M R BS SHAPE FORCED_EVERY\n");
    return 2;
  }
  const uint64_t seed = strtoull(argv[1], nullptr, 10);
  const uint32_t m = (uint32_t)strtoul(argv[2], nullptr, 10);
  const uint32_t R = (uint32_t)strtoul(argv[3], nullptr, 10);
  const uint64_t bs = strtoull(argv[4], nullptr, 10);
  const char* shape = argv[5];
  const uint32_t fevery = (uint32_t)strtoul(argv[6], nullptr, 10);
  std::mt19937_64 rng(seed);
  vector<uint32_t> szr(m), sznr(m), forced(m + 1, 0);
  for (uint32_t i = 0; i < m; i++) {
    uint32_t base;
    if (!strcmp(shape, "fixed")) base = 118 + (uint32_t)(rng() % 5);
    else if (!strcmp(shape, "varied")) base = 20 + (uint32_t)(rng() % 280);
    else if (!strcmp(shape, "small")) base = 4 + (uint32_t)(rng() % 12);
    else if (!strcmp(shape, "huge")) base = 3000 + (uint32_t)(rng() % 9000);
    else base = (rng() % 50 == 0) ? 2000 + (uint32_t)(rng() % 6000) : 30 + (uint32_t)(rng() % 150);
    sznr[i] = base;
    szr[i] = base + (uint32_t)(rng() % 24);
    if (fevery && i > 0 && rng() % fevery == 0) forced[i] = 1;
  }
  if (fevery && m > cuts::kTile) forced[cuts::kTile] = 1;  // a forced entry on a tile's first entry
  const vector<Block> ref = reference(szr, sznr, forced, bs, R);

  // tiles: functions
  const uint32_t T = (m + cuts::kTile - 1) / cuts::kTile;
  vector<TileOut> tiles(T);
  bool fallback = false;
  for (uint32_t t = 0; t < T; t++) {
    const uint64_t T0 = (uint64_t)t * cuts::kTile;
    TileOut& o = tiles[t];
    if (!stage(szr, sznr, forced, T0, bs, R, &o.te, &o.lim)) fallback = true;
    o.C = cuts::entry_points(g_tile, o.lim, o.te, T0 + o.lim == m, bs);
    if (o.C == 0 || o.C > cuts::kCand) {
      fallback = true;
      continue;
    }
    o.ftab.resize(o.C);
    for (uint32_t c = 0; c < o.C; c++) o.ftab[c] = cuts::walk(g_tile, c, o.te);
  }
  if (fallback) {
    printf("ok blocks=%zu fallback=1\n", ref.size());
    return 0;
  }
  // composition from the first tile (cut_chain), then the blocks (cut_emit)
  vector<uint32_t> entry(T), base(T);
  uint32_t e = 0, B = 0;
  for (uint32_t t = 0; t < T; t++) {
    if (e >= tiles[t].C) {
      printf("tile %u: entry %u outside its %u entry points\n", t, e, tiles[t].C);
      return 1;
    }
    entry[t] = e;
    base[t] = B;
    const uint32_t v = tiles[t].ftab[e];
    B += v >> 12;
    e = v & 0xfff;
  }
  if (e != 0) {
    printf("the chain leaves the last tile at %u\n", e);
    return 1;
  }
  if (B != ref.size()) {
    printf("blocks: tiled %u, reference %zu\n", B, ref.size());
    return 1;
  }
  vector<Block> got(B);
  for (uint32_t t = 0; t < T; t++) {
    const uint64_t T0 = (uint64_t)t * cuts::kTile;
    uint32_t te, lim;
    stage(szr, sznr, forced, T0, bs, R, &te, &lim);
    // cut_emit: the chain from the tile's entry point, the kept bytes
    vector<uint32_t> starts;
    uint32_t j = entry[t];
    while (j < te) {
      starts.push_back(j);
      j = cuts::nx_at(g_tile, j);
    }
    for (uint32_t i = 0; i < starts.size(); i++) {
      const uint32_t x = starts[i], y = cuts::nx_at(g_tile, x);
      const uint32_t id = base[t] + i;
      if (id >= B) {
        printf("tile %u: block id %u past %u\n", t, id, B);
        return 1;
      }
      got[id] = Block{(uint32_t)(T0 + x), y - x, g_bb[x], (T0 + x > 0 && forced[T0 + x]) ? 1u : 0u};
      if (g_bb[x] != (uint32_t)cuts::estimate(g_tile, x, y, R)) {
        printf("tile %u: block at %u: kept bytes %u, estimate %llu\n", t, x, g_bb[x],
               (unsigned long long)cuts::estimate(g_tile, x, y, R));
        return 1;
      }
    }
  }
  for (uint32_t i = 0; i < B; i++)
    if (got[i] != ref[i]) {
      printf("block %u: tiled (first %u, n %u, bytes %u, forced %u), reference (%u, %u, %u, %u)\n", i, got[i].first,
             got[i].n, got[i].bytes, got[i].forced, ref[i].first, ref[i].n, ref[i].bytes, ref[i].forced);
      return 1;
    }
  printf("ok blocks=%u fallback=0\n", B);
  return 0;
}

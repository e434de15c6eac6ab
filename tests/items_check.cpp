// Host checks of the batched launch's work items (beamforming_amd/csrc/
// bf_items.hpp) on random chunk lists: every (chunk, frequency, beam group,
// time tile) is covered by exactly one segment; balanced items number at
// most `grid`, cost at most the share in rounds, and cut pieces only on
// round boundaries; fixed items cover tile ranges of at most their size;
// the descriptor block's record order (descriptor_order) puts every
// segment at one slot, each item's first at the item's index.
// Prints "ok" and exits 0 when every check passes.
#include "bf_items.hpp"

#include <cstdio>
#include <map>
#include <tuple>

using namespace bf;

namespace {

int failures = 0;
void expect(bool ok, const char *what, long long a = 0, long long b = 0) {
  if (!ok && failures++ < 20) std::printf("FAIL %s (%lld, %lld)\n", what, a, b);
}

uint64_t state = 0x2545f4914f6cdd1dull;
uint64_t next() {
  state ^= state << 13;
  state ^= state >> 7;
  state ^= state << 17;
  return state;
}

int64_t tiles_of(const bf_chunk &c) { return (c.ntimes + kTileT - 1) / kTileT; }

// Every tile of every (chunk, f, bg) exactly once.
void check_cover(const std::vector<bf_chunk> &chunks, int nbg, const std::vector<BatchItem> &segs) {
  std::map<std::tuple<int, int, int>, std::vector<int>> seen;
  for (const BatchItem &s : segs) {
    expect(s.chunk >= 0 && s.chunk < (int)chunks.size(), "chunk index", s.chunk);
    if (s.chunk < 0 || s.chunk >= (int)chunks.size()) continue;
    const bf_chunk &c = chunks[s.chunk];
    expect(s.f < chunk_nfreq(c) && s.bg < nbg, "frequency or beam group", s.f, s.bg);
    expect(0 <= s.tile_begin && s.tile_begin < s.tile_end && s.tile_end <= tiles_of(c), "tile range", s.tile_begin,
           s.tile_end);
    auto &v = seen[{s.chunk, s.f, s.bg}];
    if (v.empty()) v.assign((size_t)tiles_of(c), 0);
    for (int t = s.tile_begin; t < s.tile_end && t < (int)v.size(); ++t) ++v[t];
  }
  for (int i = 0; i < (int)chunks.size(); ++i)
    for (int f = 0; f < chunk_nfreq(chunks[i]); ++f)
      for (int bg = 0; bg < nbg; ++bg) {
        auto it = seen.find({i, f, bg});
        expect(it != seen.end(), "piece missing", i, f);
        if (it == seen.end()) continue;
        for (int c : it->second) expect(c == 1, "tile covered once", c);
      }
}

// The descriptor block's record order: every segment at exactly one slot;
// item w = its first segment at slot w, then slots extra[2 w] ..
// extra[2 w + 1] - 1, in the item's order; the others' slots follow the
// first ones, item by item.
void check_order(const std::vector<int> &work) {
  std::vector<int> order, extra;
  descriptor_order(work, order, extra);
  const int nwork = (int)work.size() - 1, nseg = work.back();
  expect((int)order.size() == nseg && (int)extra.size() == 2 * nwork, "table sizes", (long long)order.size(), nseg);
  std::vector<int> seen((size_t)nseg, 0);
  for (int v : order)
    if (v >= 0 && v < nseg) ++seen[(size_t)v];
  for (int c : seen) expect(c == 1, "each segment at one slot", c);
  int expect_next = nwork;
  for (int w = 0; w < nwork; ++w) {
    expect(order[(size_t)w] == work[w], "first segment at the item's slot", w, order[(size_t)w]);
    expect(extra[2 * w] == expect_next && extra[2 * w + 1] - extra[2 * w] == work[w + 1] - work[w] - 1,
           "further segments' slots", w, extra[2 * w]);
    for (int i = 0; i < work[w + 1] - work[w] - 1 && extra[2 * w] + i < nseg; ++i)
      expect(order[(size_t)(extra[2 * w] + i)] == work[w] + 1 + i, "further segments in order", w, i);
    expect_next = extra[2 * w + 1];
  }
}

void check_balanced(const std::vector<bf_chunk> &chunks, int nbg, int grid) {
  std::vector<BatchItem> segs;
  std::vector<int> work;
  const int64_t share = balanced_items(chunks.data(), (int)chunks.size(), nbg, grid, segs, work);
  check_cover(chunks, nbg, segs);
  check_order(work);
  const int nitems = (int)work.size() - 1;
  expect(nitems <= grid, "items <= grid", nitems, grid);
  expect(work.front() == 0 && work.back() == (int)segs.size(), "item starts span the segments");
  int64_t total = 0;
  for (int w = 0; w < nitems; ++w) {
    expect(work[w] < work[w + 1], "item not empty", work[w], work[w + 1]);
    int64_t rounds = 0;
    for (int k = work[w]; k < work[w + 1]; ++k) {
      const BatchItem &s = segs[k];
      rounds += (s.tile_end - s.tile_begin + kWaves - 1) / kWaves;
      // a cut inside a piece falls on a round boundary
      expect(s.tile_begin % kWaves == 0, "segment starts on a round", s.tile_begin);
      const bool piece_end = s.tile_end == tiles_of(chunks[s.chunk]);
      expect(piece_end || s.tile_end % kWaves == 0, "segment ends on a round or the piece end", s.tile_end);
    }
    expect(rounds <= share, "item within the share", rounds, share);
    total += rounds;
  }
  int64_t want = 0;
  for (const bf_chunk &c : chunks) want += (int64_t)chunk_nfreq(c) * nbg * ((tiles_of(c) + kWaves - 1) / kWaves);
  expect(total == want, "rounds conserved", total, want);
}

void check_fixed(const std::vector<bf_chunk> &chunks, int nbg, int it) {
  std::vector<BatchItem> items;
  fixed_items(chunks.data(), (int)chunks.size(), nbg, it, items);
  check_cover(chunks, nbg, items);
  for (const BatchItem &s : items) {
    expect(s.tile_end - s.tile_begin <= it, "fixed item size", s.tile_end - s.tile_begin, it);
    expect(s.tile_begin % it == 0, "fixed item start", s.tile_begin, it);
  }
}

} // namespace

int main() {
  for (int trial = 0; trial < 3000; ++trial) {
    const int n = 1 + (int)(next() % 40);
    std::vector<bf_chunk> chunks((size_t)n);
    for (bf_chunk &c : chunks) {
      const int kind = (int)(next() % 4);
      c.ntimes = kind == 0 ? 1 + (int64_t)(next() % 40) : kind == 1 ? 1 + (int64_t)(next() % 5000)
                                                          : 16 * (1 + (int64_t)(next() % 2048));
      c.nfreq = (int)(next() % 5); // 0 means 1
      c.channel = 0;
      c.E = nullptr;
      c.J = nullptr;
    }
    const int nbg = 1 + (int)(next() % 4);
    const int grid = 1 + (int)(next() % 600);
    check_balanced(chunks, nbg, grid);
    check_fixed(chunks, nbg, kWaves * (1 + (int)(next() % 32)));
  }
  // edge cases: one tile, one chunk on a large grid, many chunks on grid 1
  check_balanced({bf_chunk{nullptr, nullptr, 1, 0, 1}}, 1, 256);
  check_balanced(std::vector<bf_chunk>(64, bf_chunk{nullptr, nullptr, 33, 0, 2}), 3, 1);
  if (failures) {
    std::printf("%d failures\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}

// bf_items.hpp -- the work items of a batched launch (host code, also
// compiled for the CPU checks in tests/items_check.cpp): the chunks' time
// tiles per (chunk, frequency, beam group) cut into segments, and segments
// grouped into work items.
#pragma once

#include "beamform.h"
#include "bf_device.hpp" // kTileT, kWaves

#include <algorithm>
#include <cstdint>
#include <vector>

namespace bf {

// One segment: tiles tile_begin .. tile_end - 1 of frequency f of a chunk,
// beam group bg.
struct BatchItem {
  int32_t chunk;
  uint16_t f, bg; // frequency within the chunk, beam group
  int32_t tile_begin, tile_end;
};

inline int chunk_nfreq(const bf_chunk &k) { return k.nfreq == 0 ? 1 : k.nfreq; }

// A segment as the batch kernel reads it: self-contained, 48 bytes (three
// 16-byte pieces), so that a workgroup loads all segments of its work item in
// one round trip.
struct SegRec {
  const uint8_t *E;  // E byte of (t = 0, d = 0, p = 0) of the segment's frequency
  uint8_t *J;        // J of the chunk
  int64_t T;         // time samples of the chunk
  int32_t channel;   // channel-table row of the segment's frequency
  int32_t nf;        // frequencies of the chunk
  int32_t f_bg;      // frequency within the chunk | beam group << 16
  int32_t tile_begin, tile_end;
  int32_t reserved;
};
static_assert(sizeof(SegRec) == 48, "three 16-byte pieces");

inline SegRec seg_record(const bf_chunk &k, const BatchItem &s, int64_t D) {
  SegRec r;
  r.E = k.E + (int64_t)s.f * D * 2;
  r.J = k.J;
  r.T = k.ntimes;
  r.channel = k.channel + s.f;
  r.nf = chunk_nfreq(k);
  r.f_bg = (int32_t)s.f | ((int32_t)s.bg << 16);
  r.tile_begin = s.tile_begin;
  r.tile_end = s.tile_end;
  r.reserved = 0;
  return r;
}

// Order of the segment records in a batch's descriptor block: item w's
// first segment at slot w, then the other segments of every item, item by
// item; extra[2 w] .. extra[2 w + 1] - 1 are the slots of item w's others.
// order[slot] is the segment (index into the segment list) at that slot.
// work holds the item starts (item w = segments work[w] .. work[w + 1] - 1).
inline void descriptor_order(const std::vector<int> &work, std::vector<int> &order, std::vector<int> &extra) {
  const int nwork = (int)work.size() - 1;
  order.assign(work.empty() ? 0 : (size_t)work.back(), -1);
  extra.assign((size_t)2 * std::max(nwork, 0), 0);
  int next = nwork;
  for (int w = 0; w < nwork; ++w) {
    order[w] = work[w];
    extra[2 * w] = next;
    for (int s = work[w] + 1; s < work[w + 1]; ++s) order[next++] = s;
    extra[2 * w + 1] = next;
  }
}

// Work items with a fixed size of `it` tiles, in chunk order.
inline void fixed_items(const bf_chunk *chunks, int n, int nbg, int it, std::vector<BatchItem> &items) {
  for (int i = 0; i < n; ++i) {
    const int64_t nt = (chunks[i].ntimes + kTileT - 1) / kTileT;
    const int nf = chunk_nfreq(chunks[i]);
    for (int f = 0; f < nf; ++f)
      for (int bg = 0; bg < nbg; ++bg)
        for (int64_t t = 0; t < nt; t += it)
          items.push_back(BatchItem{i, (uint16_t)f, (uint16_t)bg, (int32_t)t, (int32_t)std::min(nt, t + it)});
  }
}

// Balanced work items. A workgroup runs a segment of m tiles in
// ceil(m / 8) rounds (one tile per wave per round), so the cost of a
// (chunk, frequency, beam group) piece is counted in rounds, its last one
// possibly partial. Each segment after an item's first adds the cost of a
// segment start (the E pipeline refilled, a barrier, the weights restaged
// where the channel changes), about a round. The pieces are laid end
// to end and cut into items of at most `cap` in cost, the smallest cap (in
// quarter rounds, from ceil(total rounds / grid) rounds up) that needs at
// most `grid` items; a cut inside a piece falls on a round boundary, and an
// item that crosses a piece boundary becomes several segments. Every item
// then costs at most the cap, the E pipeline restarts only at segment
// starts, and the weights are restaged only where an item crosses into
// another channel. Returns the largest number of rounds an item may hold.
constexpr int64_t kSegmentStartQ = 4; // cost of a segment start, in quarter rounds

inline int64_t balanced_items(const bf_chunk *chunks, int n, int nbg, int grid, std::vector<BatchItem> &segs,
                       std::vector<int> &work) {
  int64_t total_rounds = 0;
  for (int i = 0; i < n; ++i) {
    const int64_t nt = (chunks[i].ntimes + kTileT - 1) / kTileT;
    total_rounds += (int64_t)chunk_nfreq(chunks[i]) * nbg * ((nt + kWaves - 1) / kWaves);
  }
  const size_t segs0 = segs.size(), work0 = work.size();
  // items of cost at most capq quarter rounds; returns their number
  auto fill = [&](int64_t capq) -> int64_t {
    segs.resize(segs0);
    work.resize(work0);
    work.push_back((int)segs.size());
    int64_t cost = 0, nitems = 0;
    for (int i = 0; i < n; ++i) {
      const int64_t nt = (chunks[i].ntimes + kTileT - 1) / kTileT;
      const int nf = chunk_nfreq(chunks[i]);
      for (int f = 0; f < nf; ++f)
        for (int bg = 0; bg < nbg; ++bg)
          for (int64_t t = 0; t < nt;) {
            const bool open = (int)segs.size() > work.back(); // the item holds a segment
            const int64_t extra = open ? kSegmentStartQ : 0;
            const int64_t avail = (capq - cost - extra) / 4; // whole rounds that still fit
            if (avail <= 0) { // close the item (never empty: a fresh item has room for a round)
              work.push_back((int)segs.size());
              ++nitems;
              cost = 0;
              continue;
            }
            const int64_t rounds = std::min((nt - t + kWaves - 1) / kWaves, avail);
            const int64_t take = std::min(nt - t, rounds * kWaves);
            segs.push_back(BatchItem{i, (uint16_t)f, (uint16_t)bg, (int32_t)t, (int32_t)(t + take)});
            t += take;
            cost += 4 * rounds + extra;
          }
    }
    if ((int)segs.size() > work.back()) {
      work.push_back((int)segs.size());
      ++nitems;
    }
    return nitems;
  };
  int64_t capq = 4 * std::max<int64_t>(1, (total_rounds + grid - 1) / grid);
  while (fill(capq) > grid) capq += std::max<int64_t>(1, capq / 32);
  return capq / 4;
}

} // namespace bf

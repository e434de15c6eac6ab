// bf_ingest.hip -- E ingest: packets of time samples for all channels, in the
// reference layout E[t][f][d][p] (Elinear, arraysizes.hxx:79-89), are cut
// into fixed-length chunks per group of G consecutive channels and written
// into per-group device ring buffers, [T_chunk][G][D][P] per slot: the layout
// the batch and serving paths take (bf_chunk with nfreq = G).
//
// The corner turn is done by the copy engines: for each group, one strided
// 2-D copy per packet segment (source pitch F D P, destination pitch G D P,
// width G D P bytes), from pinned host memory or from device memory, on the
// ingest's own stream. A chunk is complete when the packets have filled
// T_chunk samples; bf_ingest_pop hands it out after making the consumer's
// stream wait for its copies, and bf_ingest_release returns its slot once
// the consumer's work on it (ordered on a stream) has finished.
#include "beamform.h"
#include "bf_plan.hpp"

#include <algorithm>
#include <deque>
#include <string>
#include <vector>

using namespace bf;

namespace {

struct Slot {
  uint8_t *E = nullptr;         // [T_chunk][G][D][P]
  hipEvent_t filled = nullptr;  // copies into the slot done
  hipEvent_t released = nullptr; // consumer done (recorded on its stream)
  bool in_use = false;          // filling, waiting to be popped, or popped and not yet released
  bool release_pending = false; // released, event may not have completed
};

struct Ready {
  int group = 0, slot = 0;
};

} // namespace

struct bf_ingest {
  bf_ingest_config cfg{};
  int ngroups = 0;
  size_t row_bytes = 0;    // D P: one channel, one time
  size_t slot_bytes = 0;   // T_chunk G D P
  hipStream_t stream = nullptr;
  std::vector<std::vector<Slot>> slots; // [group][slot]
  std::vector<int> fill_slot;           // slot being filled, per group (-1: none)
  int64_t fill_pos = 0;                 // samples already in the filling slots (same for every group)
  std::deque<Ready> ready;
};

namespace {

#define BF_HIP(call)                                                                                                   \
  do {                                                                                                                 \
    const hipError_t e_ = (call);                                                                                      \
    if (e_ != hipSuccess) return hip_fail(e_, #call);                                                                  \
  } while (0)

// A slot whose consumer has released it and whose release event completed.
bool slot_free(Slot &s) {
  if (!s.in_use) return true;
  if (s.release_pending && hipEventQuery(s.released) == hipSuccess) {
    s.in_use = false;
    s.release_pending = false;
    return true;
  }
  return false;
}

} // namespace

extern "C" {

int bf_ingest_create(const bf_ingest_config *cfg, bf_ingest **out) {
  if (!out) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (!cfg) return fail(BF_ERR_INVALID_ARGUMENT, "NULL config");
  if (cfg->nfrequencies < 1 || cfg->ndishes < 1 || cfg->npolarizations != 2 || cfg->group_channels < 1 ||
      cfg->nfrequencies % cfg->group_channels != 0 || cfg->ring_slots < 2 || cfg->chunk_ntimes < 1 ||
      cfg->reserved != 0 || cfg->reserved2 != 0)
    return fail(BF_ERR_INVALID_ARGUMENT, "bad ingest config (P must be 2, G must divide F, at least 2 ring slots, "
                                         "chunk_ntimes >= 1)");
  if (cfg->ndishes % 8 != 0) return fail(BF_ERR_UNSUPPORTED, "ndishes must be a multiple of 8 (16-byte rows)");
  bf_ingest *g = new bf_ingest();
  g->cfg = *cfg;
  g->ngroups = (int)(cfg->nfrequencies / cfg->group_channels);
  g->row_bytes = (size_t)cfg->ndishes * cfg->npolarizations;
  g->slot_bytes = (size_t)cfg->chunk_ntimes * cfg->group_channels * g->row_bytes;
  g->slots.assign(g->ngroups, std::vector<Slot>(cfg->ring_slots));
  g->fill_slot.assign(g->ngroups, -1);
  auto bail = [&](int rc) {
    bf_ingest_destroy(g);
    return rc;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking)) != hipSuccess)
    return bail(hip_fail(e, "hipStreamCreate"));
  for (auto &grp : g->slots)
    for (Slot &s : grp) {
      if ((e = hipMalloc(&s.E, g->slot_bytes)) != hipSuccess) return bail(hip_fail(e, "hipMalloc (ring slot)"));
      if ((e = hipEventCreateWithFlags(&s.filled, hipEventDisableTiming)) != hipSuccess)
        return bail(hip_fail(e, "hipEventCreate"));
      if ((e = hipEventCreateWithFlags(&s.released, hipEventDisableTiming)) != hipSuccess)
        return bail(hip_fail(e, "hipEventCreate"));
    }
  *out = g;
  return BF_OK;
}

int bf_ingest_destroy(bf_ingest *g) {
  if (!g) return BF_OK;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  for (auto &grp : g->slots)
    for (Slot &s : grp) {
      if (s.release_pending) (void)hipEventSynchronize(s.released);
      if (s.E) (void)hipFree(s.E);
      if (s.filled) (void)hipEventDestroy(s.filled);
      if (s.released) (void)hipEventDestroy(s.released);
    }
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return BF_OK;
}

int bf_ingest_push(bf_ingest *g, const uint8_t *packet, int64_t ntimes, int packet_on_host) {
  if (!g || !packet) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (ntimes < 1 || (packet_on_host != 0 && packet_on_host != 1))
    return fail(BF_ERR_INVALID_ARGUMENT, "ntimes must be >= 1 and packet_on_host 0 or 1");
  const bf_ingest_config &c = g->cfg;
  // Slots this packet needs per group: the one being filled (if any) plus one
  // for every chunk boundary it starts. All must be free before anything is
  // copied, so that a refusal has no side effects.
  // (positions fill_pos .. fill_pos + ntimes - 1 that are multiples of T_chunk)
  const int64_t new_slots = g->fill_pos == 0 ? (ntimes + c.chunk_ntimes - 1) / c.chunk_ntimes
                                             : (g->fill_pos + ntimes - 1) / c.chunk_ntimes;
  if (new_slots > c.ring_slots - (g->fill_pos > 0 ? 1 : 0))
    return fail(BF_ERR_UNSUPPORTED, "packet spans more chunks than the ring holds; push it in parts");
  std::vector<std::vector<int>> plan(g->ngroups);
  for (int grp = 0; grp < g->ngroups; ++grp) {
    int exclude = g->fill_slot[grp];
    std::vector<int> taken;
    for (int64_t k = 0; k < new_slots; ++k) {
      int found = -1;
      for (int s = 0; s < c.ring_slots; ++s) {
        if (s == exclude || std::find(taken.begin(), taken.end(), s) != taken.end()) continue;
        if (slot_free(g->slots[grp][s])) {
          found = s;
          break;
        }
      }
      if (found < 0) return BF_ERR_NOT_READY; // backpressure: the consumer holds every slot
      taken.push_back(found);
    }
    plan[grp] = taken;
  }
  // Copy: segments between chunk boundaries.
  const hipMemcpyKind kind = packet_on_host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  const size_t spitch = (size_t)c.nfrequencies * g->row_bytes;
  const size_t width = (size_t)c.group_channels * g->row_bytes;
  std::vector<size_t> next(g->ngroups, 0);
  int64_t done = 0;
  while (done < ntimes) {
    if (g->fill_pos == 0)
      for (int grp = 0; grp < g->ngroups; ++grp) {
        const int s = plan[grp][next[grp]++];
        g->fill_slot[grp] = s;
        g->slots[grp][s].in_use = true;
        g->slots[grp][s].release_pending = false;
      }
    const int64_t n = std::min(ntimes - done, c.chunk_ntimes - g->fill_pos);
    for (int grp = 0; grp < g->ngroups; ++grp) {
      Slot &s = g->slots[grp][g->fill_slot[grp]];
      BF_HIP(hipMemcpy2DAsync(s.E + (size_t)g->fill_pos * width, width,
                              packet + (size_t)done * spitch + (size_t)grp * width, spitch, width, (size_t)n, kind,
                              g->stream));
    }
    done += n;
    g->fill_pos += n;
    if (g->fill_pos == c.chunk_ntimes) {
      for (int grp = 0; grp < g->ngroups; ++grp) {
        BF_HIP(hipEventRecord(g->slots[grp][g->fill_slot[grp]].filled, g->stream));
        g->ready.push_back(Ready{grp, g->fill_slot[grp]});
        g->fill_slot[grp] = -1;
      }
      g->fill_pos = 0;
    }
  }
  return BF_OK;
}

int bf_ingest_pop(bf_ingest *g, void *stream, bf_chunk *chunk, uint64_t *slot_id) {
  if (!g || !chunk) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (g->ready.empty()) return BF_ERR_NOT_READY;
  const Ready r = g->ready.front();
  Slot &s = g->slots[r.group][r.slot];
  BF_HIP(hipStreamWaitEvent(reinterpret_cast<hipStream_t>(stream), s.filled, 0));
  g->ready.pop_front();
  chunk->E = s.E;
  chunk->J = nullptr;
  chunk->ntimes = g->cfg.chunk_ntimes;
  chunk->channel = r.group * g->cfg.group_channels;
  chunk->nfreq = g->cfg.group_channels;
  if (slot_id) *slot_id = ((uint64_t)r.group << 32) | (uint32_t)r.slot;
  return BF_OK;
}

int bf_ingest_release(bf_ingest *g, uint64_t slot_id, void *stream) {
  if (!g) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  const uint64_t grp = slot_id >> 32, k = slot_id & 0xffffffffu;
  if (grp >= (uint64_t)g->ngroups || k >= (uint64_t)g->cfg.ring_slots || !g->slots[grp][k].in_use ||
      g->slots[grp][k].release_pending || (int)k == g->fill_slot[grp])
    return fail(BF_ERR_INVALID_ARGUMENT, "not a popped slot");
  for (const Ready &r : g->ready)
    if (r.group == (int)grp && r.slot == (int)k) return fail(BF_ERR_INVALID_ARGUMENT, "slot not popped yet");
  Slot &s = g->slots[grp][k];
  BF_HIP(hipEventRecord(s.released, reinterpret_cast<hipStream_t>(stream)));
  s.release_pending = true;
  return BF_OK;
}

int bf_ingest_pending(const bf_ingest *g, int64_t *filled_samples, int *ready_chunks) {
  if (!g) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (filled_samples) *filled_samples = g->fill_pos;
  if (ready_chunks) *ready_chunks = (int)g->ready.size();
  return BF_OK;
}

} // extern "C"

// bf_ingest.hip -- E ingest: packets of time samples for all channels, in the
// reference layout E[t][f][d][p] (Elinear, arraysizes.hxx:79-89), are cut
// into fixed-length chunks per group of G consecutive channels and written
// into per-group device ring buffers, [T_chunk][G][D][P] per slot: the layout
// the batch and serving paths take (bf_chunk with nfreq = G).
//
// The corner turn runs on the ingest's own stream. Device-resident packets
// (16-byte aligned) go through one scatter kernel per packet segment: every
// thread moves 16-byte pieces from a packet row to the row of its group's
// slot, reading and writing whole contiguous rows, so a packet costs one
// launch whatever the number of groups. Host packets (and unaligned device
// packets) go through the copy engines: for each group, one strided 2-D copy
// per packet segment (source pitch F D P, destination pitch G D P, width
// G D P bytes). A chunk is complete when the packets have filled T_chunk
// samples; bf_ingest_pop hands it out after making the consumer's stream wait
// for its copies, and bf_ingest_release returns its slot once the consumer's
// work on it (ordered on a stream) has finished.
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

// Ring slot of each group of one scatter launch (groups grp0 .. grp0 + n - 1).
constexpr int kScatterGroups = 256;
struct SlotPick {
  uint8_t slot[kScatterGroups];
};

// Copies rows t = 0 .. n - 1 of a packet segment: group g's width bytes at
// packet + t spitch + (grp0 + g) width go to its slot at dst_off + t width.
// width and spitch are multiples of 16, the packet is 16-byte aligned.
// blockIdx.x covers the 16-byte pieces of a row's groups (row_units =
// ngroups width / 16 < 2^31), blockIdx.y strides over the rows.
__global__ void ingest_scatter_kernel(const uint8_t *__restrict__ packet, int64_t n, int64_t spitch, int64_t width,
                                      int grp0, int ngroups, uint8_t *const *__restrict__ slot_table, int ring_slots,
                                      SlotPick pick, int64_t dst_off) {
  const uint32_t wu = (uint32_t)(width / 16);
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= (uint32_t)ngroups * wu) return;
  const uint32_t g = u / wu, k = u - g * wu;
  const uint8_t *src = packet + (int64_t)(grp0 + (int)g) * width + 16 * (int64_t)k;
  uint8_t *dst = slot_table[(int64_t)(grp0 + (int)g) * ring_slots + pick.slot[g]] + dst_off + 16 * (int64_t)k;
  for (int64_t t = blockIdx.y; t < n; t += gridDim.y)
    *reinterpret_cast<uint4 *>(dst + t * width) = *reinterpret_cast<const uint4 *>(src + t * spitch);
}

} // namespace

struct bf_ingest {
  bf_ingest_config cfg{};
  int ngroups = 0;
  size_t row_bytes = 0;    // D P: one channel, one time
  size_t slot_bytes = 0;   // T_chunk G D P
  hipStream_t stream = nullptr;
  std::vector<std::vector<Slot>> slots; // [group][slot]
  uint8_t **d_slot_table = nullptr;     // device: slot base addresses, [group][slot]
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
  if (cfg->nfrequencies * cfg->ndishes * cfg->npolarizations / 16 >= (1ll << 31))
    return fail(BF_ERR_UNSUPPORTED, "packet rows of 32 GiB or more");
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
  {
    std::vector<uint8_t *> table;
    for (auto &grp : g->slots)
      for (Slot &s : grp) table.push_back(s.E);
    if ((e = hipMalloc(&g->d_slot_table, table.size() * sizeof(uint8_t *))) != hipSuccess)
      return bail(hip_fail(e, "hipMalloc (slot table)"));
    if ((e = hipMemcpy(g->d_slot_table, table.data(), table.size() * sizeof(uint8_t *), hipMemcpyHostToDevice)) !=
        hipSuccess)
      return bail(hip_fail(e, "hipMemcpy (slot table)"));
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
  if (g->d_slot_table) (void)hipFree(g->d_slot_table);
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
  // device packets: one scatter kernel per segment (widths are multiples of
  // 16 because D is a multiple of 8)
  bool scatter = !packet_on_host && ((uintptr_t)packet & 15u) == 0 && c.ring_slots <= 255;
  if (scatter) { // only memory the kernel can read for sure: device allocations
    hipPointerAttribute_t attr;
    scatter = hipPointerGetAttributes(&attr, packet) == hipSuccess &&
              (attr.type == hipMemoryTypeDevice || attr.isManaged);
    (void)hipGetLastError(); // a failed query of a host pointer is not an error here
  }
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
    if (scatter) {
      for (int grp0 = 0; grp0 < g->ngroups; grp0 += kScatterGroups) {
        const int ng = std::min(kScatterGroups, g->ngroups - grp0);
        SlotPick pick;
        for (int i = 0; i < ng; ++i) pick.slot[i] = (uint8_t)g->fill_slot[grp0 + i];
        const int64_t row_units = (int64_t)ng * (int64_t)(width / 16); // < 2^31, checked at creation
        const dim3 grid((unsigned)((row_units + 255) / 256), (unsigned)std::min<int64_t>(n, 65535));
        hipLaunchKernelGGL(ingest_scatter_kernel, grid, dim3(256), 0, g->stream, packet + (size_t)done * spitch,
                           n, (int64_t)spitch, (int64_t)width, grp0, ng, g->d_slot_table, c.ring_slots, pick,
                           (int64_t)g->fill_pos * (int64_t)width);
        BF_HIP(hipGetLastError());
      }
    } else {
      for (int grp = 0; grp < g->ngroups; ++grp) {
        Slot &s = g->slots[grp][g->fill_slot[grp]];
        BF_HIP(hipMemcpy2DAsync(s.E + (size_t)g->fill_pos * width, width,
                                packet + (size_t)done * spitch + (size_t)grp * width, spitch, width, (size_t)n, kind,
                                g->stream));
      }
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

int bf_ingest_sync(bf_ingest *g) {
  if (!g) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  BF_HIP(hipStreamSynchronize(g->stream));
  return BF_OK;
}

int bf_ingest_pending(const bf_ingest *g, int64_t *filled_samples, int *ready_chunks) {
  if (!g) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (filled_samples) *filled_samples = g->fill_pos;
  if (ready_chunks) *ready_chunks = (int)g->ready.size();
  return BF_OK;
}

} // extern "C"

// bf_ingest.hip -- E ingest: packets of time samples for all channels, in the
// reference layout E[t][f][d][p] (Elinear, arraysizes.hxx:79-89), are cut
// into fixed-length chunks per group of G consecutive channels and written
// into per-group device ring buffers, [T_chunk][G][D][P] per slot: the layout
// the batch and serving paths take (bf_chunk with nfreq = G).
//
// The corner turn runs on the ingest's own stream. Device-resident packets
// (16-byte aligned) are copied by a scatter kernel that moves whole packet
// rows (all groups of one time sample per workgroup) to their groups' slot
// rows. A push only records its packet segments; one launch copies every
// recorded segment when a chunk completes (so its completion event follows
// its copies), when bf_ingest_sync is called, or when kMaxPending segments are
// waiting, so that a stream of small packets costs one launch per chunk, not
// one per packet. Host packets (and unaligned device packets) go through the
// copy engines at once: for each group, one strided 2-D copy per packet
// segment (source pitch F D P, destination pitch G D P, width G D P bytes). A
// chunk is complete when the packets have filled T_chunk samples;
// bf_ingest_pop hands it out after making the consumer's stream wait for its
// copies, and bf_ingest_release returns its slot once the consumer's work on
// it (ordered on a stream) has finished.
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
  hipEvent_t released = nullptr; // consumer done (recorded on its stream)
  bool in_use = false;          // filling, waiting to be popped, or popped and not yet released
  bool release_pending = false; // released, event may not have completed
  uint64_t release_seq = 0;     // order of the release calls (1, 2, ...)
};

struct Ready {
  int group = 0, slot = 0;
  uint64_t seq = 0; // chunk time (0, 1, ...) the slot holds
};

// Ring slot of each group of one scatter launch (groups grp0 .. grp0 + n - 1).
constexpr int kScatterGroups = 256;
struct SlotPick {
  uint8_t slot[kScatterGroups];
};

// Packet segments copied by one launch: segment k holds n[k] rows starting
// at src[k] (pitch spitch) that go to rows dst_row[k] .. of the groups' fill
// slots; rows_before[k] = n[0] + .. + n[k - 1]. All segments of a launch fill
// the same slots (a launch never spans a chunk boundary).
constexpr int kMaxPending = 32;
struct ScatterBatch {
  const uint8_t *src[kMaxPending];
  int64_t dst_row[kMaxPending];
  int64_t rows_before[kMaxPending + 1];
  int32_t nseg;
};

// Workgroup r copies row r of the launch (segment k, row r - rows_before[k]):
// its ngroups * width bytes, contiguous in the packet, go to width bytes in
// each group's slot. width and spitch are multiples of 16 and the packets are
// 16-byte aligned; each thread moves kPieces 16-byte pieces per pass, all
// loads issued before the stores. The stores are non-temporal: the slots are
// read by a later launch, not from this one's caches (a 256 MiB corner turn
// ran at 6.9 TB/s of read plus write this way, 5.4 TB/s with plain stores).
typedef unsigned v4u __attribute__((ext_vector_type(4)));
constexpr int kScatterThreads = 256;
constexpr int kPieces = 4;
__global__ __launch_bounds__(kScatterThreads) void ingest_scatter_kernel(const ScatterBatch b, int64_t spitch,
                                                                          int64_t width, int grp0, int ngroups,
                                                                          uint8_t *const *__restrict__ slot_table,
                                                                          int ring_slots, SlotPick pick) {
  const int64_t r = blockIdx.x;
  int k = 0;
  while (k + 1 < b.nseg && b.rows_before[k + 1] <= r) ++k; // uniform, at most kMaxPending steps
  const int64_t t = r - b.rows_before[k];
  const v4u *src = reinterpret_cast<const v4u *>(b.src[k] + t * spitch + (int64_t)grp0 * width);
  const int64_t dst_off = (b.dst_row[k] + t) * width;
  const uint32_t wu = (uint32_t)(width / 16);
  const uint32_t units = (uint32_t)ngroups * wu;
  // The row's destination in every group's slot, looked up once per
  // workgroup: a store then waits on no memory read for its address.
  __shared__ __attribute__((address_space(1))) v4u *s_dst[kScatterGroups];
  for (int i = threadIdx.x; i < ngroups; i += kScatterThreads)
    s_dst[i] = (__attribute__((address_space(1))) v4u *)(slot_table[(int64_t)(grp0 + i) * ring_slots + pick.slot[i]] +
                                                          dst_off);
  __syncthreads();
  for (uint32_t u0 = threadIdx.x; u0 < units; u0 += kScatterThreads * kPieces) {
    v4u v[kPieces];
#pragma unroll
    for (int i = 0; i < kPieces; ++i) {
      const uint32_t u = u0 + (uint32_t)(i * kScatterThreads);
      if (u < units) v[i] = src[u];
    }
#pragma unroll
    for (int i = 0; i < kPieces; ++i) {
      const uint32_t u = u0 + (uint32_t)(i * kScatterThreads);
      if (u < units) {
        const uint32_t g = u / wu, q = u - g * wu;
        __builtin_nontemporal_store(v[i], s_dst[g] + q);
      }
    }
  }
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
  // One completion event per chunk time, shared by all groups (one launch,
  // or one set of 2-D copies, fills every group's slot of a chunk time);
  // chunk time k uses done[k % ring_slots]. Chunk time k + ring_slots can
  // only start once every group's slot of chunk time k is free again, i.e.
  // popped (every wait on done[k % ring_slots] issued) and released.
  std::vector<hipEvent_t> done;
  uint64_t chunk_seq = 0;               // chunk time being filled
  hipStream_t last_wait_stream = nullptr; // the stream that last waited, and for which chunk time
  uint64_t last_wait_seq = ~0ull;
  ScatterBatch pending{};               // device packet segments recorded, not yet copied
  int64_t pending_rows = 0;
  // Device allocation of the last device packet ([dev_base, dev_base +
  // dev_size)): later packets inside it skip the pointer-attribute query.
  uintptr_t dev_base = 0;
  size_t dev_size = 0;
  uint64_t release_count = 0; // bf_ingest_release calls so far
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

// Launches the copies of the recorded device packet segments (into the
// groups' current fill slots) and clears the record.
int flush_pending(bf_ingest *g) {
  ScatterBatch &b = g->pending;
  if (b.nseg == 0) return BF_OK;
  const bf_ingest_config &c = g->cfg;
  const int64_t spitch = c.nfrequencies * (int64_t)g->row_bytes;
  const int64_t width = c.group_channels * (int64_t)g->row_bytes;
  b.rows_before[b.nseg] = g->pending_rows;
  for (int grp0 = 0; grp0 < g->ngroups; grp0 += kScatterGroups) {
    const int ng = std::min(kScatterGroups, g->ngroups - grp0);
    SlotPick pick;
    for (int i = 0; i < ng; ++i) pick.slot[i] = (uint8_t)g->fill_slot[grp0 + i];
    hipLaunchKernelGGL(ingest_scatter_kernel, dim3((unsigned)g->pending_rows), dim3(kScatterThreads), 0, g->stream, b,
                       spitch, width, grp0, ng, g->d_slot_table, c.ring_slots, pick);
    BF_HIP(hipGetLastError());
  }
  b.nseg = 0;
  g->pending_rows = 0;
  return BF_OK;
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
      if ((e = hipEventCreateWithFlags(&s.released, hipEventDisableTiming)) != hipSuccess)
        return bail(hip_fail(e, "hipEventCreate"));
    }
  g->done.assign((size_t)cfg->ring_slots, nullptr);
  for (hipEvent_t &ev : g->done)
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return bail(hip_fail(e, "hipEventCreate"));
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
      if (s.released) (void)hipEventDestroy(s.released);
    }
  for (hipEvent_t ev : g->done)
    if (ev) (void)hipEventDestroy(ev);
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
    const uintptr_t p = (uintptr_t)packet;
    const size_t bytes = (size_t)ntimes * (size_t)c.nfrequencies * g->row_bytes;
    if (!(g->dev_size && p >= g->dev_base && p + bytes <= g->dev_base + g->dev_size)) {
      hipPointerAttribute_t attr;
      scatter = hipPointerGetAttributes(&attr, packet) == hipSuccess &&
                (attr.type == hipMemoryTypeDevice || attr.isManaged);
      (void)hipGetLastError(); // a failed query of a host pointer is not an error here
      void *base = nullptr;
      size_t size = 0;
      if (scatter && hipMemGetAddressRange(&base, &size, const_cast<uint8_t *>(packet)) == hipSuccess &&
          p + bytes <= (uintptr_t)base + size) {
        g->dev_base = (uintptr_t)base;
        g->dev_size = size;
      } else {
        (void)hipGetLastError();
        g->dev_size = 0;
      }
    }
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
      // recorded; copied with the other segments of this chunk in one launch
      if (g->pending.nseg == kMaxPending || g->pending_rows + n > (int64_t)0x7fffffff)
        if (int rc = flush_pending(g)) return rc;
      ScatterBatch &b = g->pending;
      b.src[b.nseg] = packet + (size_t)done * spitch;
      b.dst_row[b.nseg] = g->fill_pos;
      b.rows_before[b.nseg] = g->pending_rows;
      ++b.nseg;
      g->pending_rows += n;
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
      if (int rc = flush_pending(g)) return rc; // the chunk's copies, before its completion event
      BF_HIP(hipEventRecord(g->done[g->chunk_seq % g->done.size()], g->stream));
      for (int grp = 0; grp < g->ngroups; ++grp) {
        g->ready.push_back(Ready{grp, g->fill_slot[grp], g->chunk_seq});
        g->fill_slot[grp] = -1;
      }
      ++g->chunk_seq;
      g->fill_pos = 0;
    }
  }
  return BF_OK;
}

int bf_ingest_push_many(bf_ingest *g, const uint8_t *const *packets, const int64_t *ntimes, int n,
                        int packet_on_host, int *pushed) {
  if (pushed) *pushed = 0;
  if (!g || (n > 0 && (!packets || !ntimes)) || n < 0) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument or n < 0");
  for (int i = 0; i < n; ++i) {
    const int rc = bf_ingest_push(g, packets[i], ntimes[i], packet_on_host);
    if (rc) return rc;
    if (pushed) *pushed = i + 1;
  }
  return BF_OK;
}

int bf_ingest_pop(bf_ingest *g, void *stream, bf_chunk *chunk, uint64_t *slot_id) {
  if (!g || !chunk) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (g->ready.empty()) return BF_ERR_NOT_READY;
  const Ready r = g->ready.front();
  Slot &s = g->slots[r.group][r.slot];
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (st != g->last_wait_stream || r.seq != g->last_wait_seq) { // this stream has not waited for it yet
    BF_HIP(hipStreamWaitEvent(st, g->done[r.seq % g->done.size()], 0));
    g->last_wait_stream = st;
    g->last_wait_seq = r.seq;
  }
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
  s.release_seq = ++g->release_count;
  return BF_OK;
}

int bf_ingest_wait_release(bf_ingest *g) {
  if (!g) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  // the oldest release whose event may not have completed yet
  Slot *oldest = nullptr;
  for (auto &grp : g->slots)
    for (Slot &s : grp)
      if (s.in_use && s.release_pending && (!oldest || s.release_seq < oldest->release_seq)) oldest = &s;
  if (!oldest) return BF_ERR_NOT_READY; // nothing released: only the consumer can free a slot
  BF_HIP(hipEventSynchronize(oldest->released));
  oldest->in_use = false;
  oldest->release_pending = false;
  return BF_OK;
}

int bf_ingest_sync(bf_ingest *g) {
  if (!g) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (int rc = flush_pending(g)) return rc;
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

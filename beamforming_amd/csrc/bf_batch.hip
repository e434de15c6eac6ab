// bf_batch.hip -- one-channel launches, the batched persistent launch (its
// descriptors in pinned host memory) with hipGraph replay, and the
// autotuners.
#include "beamform.h"
#include "bf_items.hpp"
#include "bf_plan.hpp"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

using namespace bf;

#define BF_HIP(call)                                                                                                   \
  do {                                                                                                                 \
    const hipError_t e_ = (call);                                                                                      \
    if (e_ != hipSuccess) return hip_fail(e_, #call);                                                                  \
  } while (0)

struct bf_batch {
  const bf_plan *plan = nullptr;
  int max_chunks = 0;
  int item_capacity = 0;
  int item_tiles = 0; // 0 = automatic
  int last_item_tiles = 0;
  int last_nitems = 0;
  bool last_queue = false; // the last prepared run takes its items from the queue
  int last_segments = 0;
  int grid = 0;
  int grid_narrow = 0; // with the plan's one-beam-tile configuration (0: none)
  // Two descriptor blocks in pinned host memory, used by alternate runs and
  // read by the kernel in place: [work item count, queue base, mode | pad to
  // 16][per work item, the range of its further segment records: 2 x
  // item_capacity ints, padded to 16][segment records: item_capacity; item
  // w's first at index w, then the others of every item]. Both tables hold
  // at least one entry per workgroup of a launch (each workgroup reads its
  // own item's entries before it knows the item count). A run rewrites its
  // block only after the launch that last read it (two runs before) has
  // finished.
  struct Block {
    uint8_t *h = nullptr; // host address
    uint8_t *d = nullptr; // the same memory as the kernel addresses it
    unsigned *head = nullptr; // device: the queue counter of launches reading this block
    uint32_t base = 0;        // its value when the next launch starts (it only grows)
    hipEvent_t launch_done = nullptr;
    bool launch_recorded = false;
    hipGraph_t graph = nullptr; // captured launch reading this block
    hipGraphExec_t exec = nullptr;
    bool narrow = false;        // the run written into the block takes one beam tile per workgroup
    bool graph_narrow = false;  // the configuration the graph was captured with
  } blk[2];
  int next_block = 0;
  unsigned *d_heads = nullptr; // the two queue counters, 256 bytes apart
  size_t off_extra = 16, off_items = 0, desc_bytes = 0;
  hipStream_t capture_stream = nullptr;
  uint64_t plan_version = 0; // the plan's layout_version the grid and the graphs were made for
};

namespace {

int check_channel_args(const bf_plan *p, int channel, int nfreq, int64_t ntimes, const void *J, const void *E) {
  if (!p || !J || !E) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (nfreq < 1) return fail(BF_ERR_INVALID_ARGUMENT, "nfreq must be >= 1");
  if (channel < 0 || (int64_t)channel + nfreq > p->shape.nfrequencies)
    return fail(BF_ERR_INVALID_ARGUMENT, "channels out of range");
  if (ntimes < 1) return fail(BF_ERR_INVALID_ARGUMENT, "ntimes must be >= 1");
  if (ntimes > kMaxTimes) return fail(BF_ERR_UNSUPPORTED, "ntimes too large");
  return BF_OK;
}

void free_graphs(bf_batch *b) {
  for (bf_batch::Block &k : b->blk) {
    if (k.exec) (void)hipGraphExecDestroy(k.exec);
    if (k.graph) (void)hipGraphDestroy(k.graph);
    k.exec = nullptr;
    k.graph = nullptr;
  }
}

// Waits until no launch reads either descriptor block.
int wait_launches(bf_batch *b) {
  for (bf_batch::Block &k : b->blk)
    if (k.launch_recorded) BF_HIP(hipEventSynchronize(k.launch_done));
  return BF_OK;
}

int ensure_item_capacity(bf_batch *b, int n) {
  if (n <= b->item_capacity) return BF_OK;
  int cap = std::max(n, 2 * b->item_capacity);
  if (int rc = wait_launches(b)) return rc;
  free_graphs(b); // the captured launches refer to the old blocks
  for (bf_batch::Block &k : b->blk) {
    if (k.h) (void)hipHostFree(k.h);
    k.h = k.d = nullptr;
  }
  b->item_capacity = 0;
  auto up16 = [](size_t x) { return (x + 15) / 16 * 16; };
  b->off_extra = 16;
  b->off_items = b->off_extra + up16((size_t)2 * cap * sizeof(int));
  b->desc_bytes = b->off_items + (size_t)cap * sizeof(SegRec);
  for (bf_batch::Block &k : b->blk) {
    // coherent: the device reads what the host wrote before the launch, uncached
    BF_HIP(hipHostMalloc(&k.h, b->desc_bytes, hipHostMallocCoherent));
    BF_HIP(hipHostGetDevicePointer(reinterpret_cast<void **>(&k.d), k.h, 0));
  }
  b->item_capacity = cap;
  return BF_OK;
}


// Launch geometry for the plan's current configuration: the batch kernel's
// LDS limit (the instantiation's largest layout, so plans of either layout
// can share it) and one wave of resident workgroups. A captured graph of an
// earlier configuration is dropped (it holds that configuration's kernel and
// weight table addresses).
int follow_plan(bf_batch *b) {
  const bf_plan *p = b->plan;
  if (b->grid > 0 && b->plan_version == p->layout_version) return BF_OK;
  if (!p->use_mfma)
    return fail(BF_ERR_UNSUPPORTED, "batching needs the MFMA path (P = 2, D % 8 == 0, D <= 1024, gains in range)");
  const MfmaConfig &c = p->cfg;
  if (c.nbg > 65535) return fail(BF_ERR_UNSUPPORTED, "too many beam groups for a batch work item");
  if (int rc = wait_launches(b)) return rc;
  free_graphs(b);
  BF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(c.batch_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)mfma_lds_bytes(c.nt, c.ks)));
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(c.batch_kernel), kThreads,
                                                   c.lds_bytes) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  b->grid = p->num_cus * per_cu;
  b->grid_narrow = 0;
  if (p->has_narrow) {
    const MfmaConfig &n = p->cfg_narrow;
    BF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(n.batch_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)mfma_lds_bytes(n.nt, n.ks)));
    int per_cu_n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_n, reinterpret_cast<const void *>(n.batch_kernel),
                                                     kThreads, n.lds_bytes) != hipSuccess ||
        per_cu_n < 1)
      per_cu_n = 1;
    b->grid_narrow = p->num_cus * per_cu_n;
  }
  b->plan_version = p->layout_version;
  return BF_OK;
}

// The configuration, grid and beam groups of a run in block j.
const MfmaConfig &run_config(const bf_batch *b, int j) {
  return b->blk[j].narrow ? b->plan->cfg_narrow : b->plan->cfg;
}
int run_grid(const bf_batch *b, int j) { return b->blk[j].narrow ? b->grid_narrow : b->grid; }

// Validates the chunk list and writes it and the work items (longest first)
// into the next descriptor block, whose index goes to *block.
int prepare(bf_batch *b, const bf_chunk *chunks, int n, int *block) {
  const bf_plan *p = b->plan;
  if (int rc = follow_plan(b)) return rc;
  if (n < 0 || n > b->max_chunks) return fail(BF_ERR_INVALID_ARGUMENT, "nchunks outside [0, max_chunks]");
  if (n > 0 && !chunks) return fail(BF_ERR_INVALID_ARGUMENT, "chunks is NULL");
  const MfmaConfig &c = p->cfg;
  for (int i = 0; i < n; ++i) {
    const bf_chunk &k = chunks[i];
    const int64_t nf = chunk_nfreq(k);
    if (!k.E || !k.J || k.ntimes < 1 || nf < 1 || nf > 65535 || k.channel < 0 ||
        k.channel + nf > p->shape.nfrequencies)
      return fail(BF_ERR_INVALID_ARGUMENT, "bad chunk " + std::to_string(i) +
                                               " (NULL buffer, ntimes < 1, or channels outside the table)");
  }
  for (int i = 0; i < n; ++i)
    if (chunks[i].ntimes > kMaxTimes) return fail(BF_ERR_UNSUPPORTED, "chunk too long");
  std::vector<BatchItem> items; // segments
  std::vector<int> work;        // work item starts
  int it = 0;
  bool narrow = false;
  if (b->item_tiles > 0) {
    fixed_items(chunks, n, c.nbg, b->item_tiles, items);
    // longest first, one segment per work item
    std::stable_sort(items.begin(), items.end(), [](const BatchItem &x, const BatchItem &y) {
      return x.tile_end - x.tile_begin > y.tile_end - y.tile_begin;
    });
    for (int i = 0; i <= (int)items.size(); ++i) work.push_back(i);
    it = b->item_tiles;
  } else {
    // Balanced items, one per workgroup. A batch too small to give every
    // workgroup an item runs one beam tile per workgroup instead (the
    // plan's one-beam-tile configuration over the same weight tables), with
    // NT times the items.
    int64_t rounds = 0;
    for (int i = 0; i < n; ++i)
      rounds += (int64_t)chunk_nfreq(chunks[i]) * c.nbg * (((chunks[i].ntimes + kTileT - 1) / kTileT + kWaves - 1) / kWaves);
    narrow = b->grid_narrow > 0 && rounds < b->grid;
    const MfmaConfig &cu = narrow ? p->cfg_narrow : c;
    it = (int)(balanced_items(chunks, n, cu.nbg, narrow ? b->grid_narrow : b->grid, items, work) * kWaves);
  }
  const int nwork = (int)work.size() - 1;
  const int grid = narrow ? b->grid_narrow : b->grid;
  if (int rc = ensure_item_capacity(b, std::max<int>({1, (int)items.size(), grid}))) return rc;
  const int j = b->next_block;
  bf_batch::Block &k = b->blk[j];
  if (k.launch_recorded) BF_HIP(hipEventSynchronize(k.launch_done)); // the run two before still reads it
  // item w's first segment record at w, its others after all first ones
  std::vector<int> order, extra;
  descriptor_order(work, order, extra);
  SegRec *hs = reinterpret_cast<SegRec *>(k.h + b->off_items);
  for (size_t i = 0; i < order.size(); ++i)
    hs[i] = seg_record(chunks[items[order[i]].chunk], items[order[i]], p->shape.ndishes);
  std::copy(extra.begin(), extra.end(), reinterpret_cast<int *>(k.h + b->off_extra));
  reinterpret_cast<int *>(k.h)[0] = nwork;
  reinterpret_cast<uint32_t *>(k.h)[1] = k.base;
  reinterpret_cast<int *>(k.h)[2] = b->item_tiles > 0 ? 1 : 0; // balanced items: one per workgroup, no queue
  b->last_queue = b->item_tiles > 0;
  k.narrow = narrow;
  std::atomic_thread_fence(std::memory_order_seq_cst); // the block is written before the launch is enqueued
  b->next_block = j ^ 1;
  *block = j;
  b->last_item_tiles = it;
  b->last_nitems = nwork;
  b->last_segments = (int)items.size();
  return BF_OK;
}

BatchParams batch_params(const bf_batch *b, int j) {
  const bf_plan *p = b->plan;
  const uint8_t *d = b->blk[j].d;
  BatchParams a;
  a.segs = reinterpret_cast<const SegRec *>(d + b->off_items);
  a.extra = reinterpret_cast<const int *>(d + b->off_extra);
  a.nitems = reinterpret_cast<const int *>(d);
  a.head = b->blk[j].head;
  a.W = p->W;
  a.corr = p->corr;
  a.G = p->Gp;
  a.NTP = p->cfg.ntp; // the one-beam-tile configuration shares the tables
  a.D = (int32_t)p->shape.ndishes;
  a.B = (int32_t)p->shape.nbeams;
  a.jlayout = p->shape.j_layout;
  a.stage_j = run_config(b, j).stage_j ? 1 : 0;
  return a;
}

// The launch itself: the persistent kernel reading descriptor block j.
int enqueue_launch(bf_batch *b, int j, hipStream_t st) {
  const MfmaConfig &c = run_config(b, j);
  BatchParams a = batch_params(b, j);
  void *args[] = {&a};
  BF_HIP(hipLaunchKernel(reinterpret_cast<const void *>(c.batch_kernel), dim3(run_grid(b, j)), dim3(kThreads), args,
                         c.lds_bytes, st));
  return BF_OK;
}

// After a launch on block j was enqueued: its queue counter will have grown
// by the items fetched from it (those beyond the grid; every workgroup runs
// its own item first) plus one failed fetch per workgroup.
int record_launch(bf_batch *b, int j, hipStream_t st) {
  const int grid = run_grid(b, j);
  if (b->last_queue) b->blk[j].base += (uint32_t)(std::max(0, b->last_nitems - grid) + grid);
  BF_HIP(hipEventRecord(b->blk[j].launch_done, st));
  b->blk[j].launch_recorded = true;
  return BF_OK;
}

int time_launches(hipStream_t st, int iters, float *ms, const std::function<int()> &launch) {
  hipEvent_t e0, e1;
  BF_HIP(hipEventCreate(&e0));
  BF_HIP(hipEventCreate(&e1));
  int rc = launch(); // warm-up
  if (!rc) {
    const hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = hip_fail(e, "warm-up launch");
  }
  if (!rc && hipEventRecord(e0, st) != hipSuccess) rc = BF_ERR_DEVICE;
  for (int i = 0; !rc && i < iters; ++i) rc = launch();
  if (!rc && hipEventRecord(e1, st) != hipSuccess) rc = BF_ERR_DEVICE;
  if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = BF_ERR_DEVICE;
  float t = 0;
  if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = BF_ERR_DEVICE;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = iters > 0 ? t / iters : t;
  return rc;
}

} // namespace

extern "C" {

int bf_form_beams_channels(const bf_plan *p, int channel, int nfreq, int64_t ntimes, uint8_t *J, const uint8_t *E,
                           void *stream) {
  if (int rc = check_channel_args(p, channel, nfreq, ntimes, J, E)) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const bf_shape &s = p->shape;
  if (p->use_mfma) { // any alignment of E and J
    MfmaConfig c;
    if (!pick_config(p, ntimes, nfreq, &c)) return fail(BF_ERR_UNSUPPORTED, "ntimes too large");
    return launch_mfma(p, c, J, E, ntimes, nfreq, channel, st);
  }
  const int64_t total = s.nbeams * nfreq * s.npolarizations * ntimes;
  const int blocks = (int)std::min<int64_t>((total + 255) / 256, 1 << 20);
  hipLaunchKernelGGL(form_beams_generic_kernel, dim3(blocks), dim3(256), 0, st, E, J,
                     p->A_dev + (int64_t)channel * s.nbeams * s.ndishes, p->G_dev + (int64_t)channel * s.nbeams,
                     ntimes, nfreq, (int)s.ndishes, (int)s.npolarizations, (int)s.nbeams, s.j_layout);
  BF_HIP(hipGetLastError());
  return BF_OK;
}

int bf_form_beams_channel(const bf_plan *p, int channel, int64_t ntimes, uint8_t *J, const uint8_t *E, void *stream) {
  return bf_form_beams_channels(p, channel, 1, ntimes, J, E, stream);
}

int bf_batch_create(const bf_plan *p, int max_chunks, bf_batch **out) {
  if (!out) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (!p || max_chunks < 1) return fail(BF_ERR_INVALID_ARGUMENT, "NULL plan or max_chunks < 1");
  bf_batch *b = new bf_batch();
  b->plan = p;
  b->max_chunks = max_chunks;
  auto bail = [&](int rc) {
    bf_batch_destroy(b);
    return rc;
  };
  if (int rc = follow_plan(b)) return bail(rc);
  hipError_t e;
  for (bf_batch::Block &k : b->blk)
    if ((e = hipEventCreateWithFlags(&k.launch_done, hipEventDisableTiming)) != hipSuccess)
      return bail(hip_fail(e, "hipEventCreate"));
  if ((e = hipMalloc(&b->d_heads, 512)) != hipSuccess) return bail(hip_fail(e, "hipMalloc (queue counters)"));
  if ((e = hipMemset(b->d_heads, 0, 512)) != hipSuccess) return bail(hip_fail(e, "hipMemset (queue counters)"));
  b->blk[0].head = b->d_heads;
  b->blk[1].head = b->d_heads + 64;
  if ((e = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking)) != hipSuccess)
    return bail(hip_fail(e, "hipStreamCreate"));
  if (int rc = ensure_item_capacity(b, 4 * b->grid)) return bail(rc);
  *out = b;
  return BF_OK;
}

int bf_batch_destroy(bf_batch *b) {
  if (!b) return BF_OK;
  (void)wait_launches(b); // the launches read the descriptor blocks
  free_graphs(b);
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
  for (bf_batch::Block &k : b->blk) {
    if (k.launch_done) (void)hipEventDestroy(k.launch_done);
    if (k.h) (void)hipHostFree(k.h);
  }
  if (b->d_heads) (void)hipFree(b->d_heads);
  delete b;
  return BF_OK;
}

int bf_batch_set_item_tiles(bf_batch *b, int item_tiles) {
  if (!b || item_tiles < 0) return fail(BF_ERR_INVALID_ARGUMENT, "bad argument");
  b->item_tiles = item_tiles == 0 ? 0 : std::max(kWaves, (item_tiles + kWaves - 1) / kWaves * kWaves);
  return BF_OK;
}

int bf_batch_run(bf_batch *b, const bf_chunk *chunks, int nchunks, void *stream) {
  if (!b) return fail(BF_ERR_INVALID_ARGUMENT, "NULL batch");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  int j = 0;
  if (int rc = prepare(b, chunks, nchunks, &j)) return rc;
  if (int rc = enqueue_launch(b, j, st)) return rc;
  return record_launch(b, j, st);
}

int bf_batch_run_graph(bf_batch *b, const bf_chunk *chunks, int nchunks, void *stream) {
  if (!b) return fail(BF_ERR_INVALID_ARGUMENT, "NULL batch");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  int j = 0;
  if (int rc = prepare(b, chunks, nchunks, &j)) return rc;
  bf_batch::Block &k = b->blk[j];
  if (k.exec && k.graph_narrow != k.narrow) { // captured with the other configuration
    (void)hipGraphExecDestroy(k.exec);
    (void)hipGraphDestroy(k.graph);
    k.exec = nullptr;
    k.graph = nullptr;
  }
  if (!k.exec) {
    // Capture once per block: the launch reads chunks, items and their count
    // from the block, so the graph stays valid for any chunk list.
    BF_HIP(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_launch(b, j, b->capture_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->capture_stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    k.graph = g;
    k.graph_narrow = k.narrow;
    BF_HIP(hipGraphInstantiate(&k.exec, k.graph, nullptr, nullptr, 0));
  }
  BF_HIP(hipGraphLaunch(k.exec, st));
  return record_launch(b, j, st);
}

int bf_batch_describe(const bf_batch *b, char *buf, size_t len) {
  if (!b || !buf || len == 0) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  const int last = b->next_block ^ 1; // the block of the last run
  const MfmaConfig &c = run_config(b, last);
  std::snprintf(buf, len,
                "persistent %s batch: nt=%d ks=%d nbg=%d stage_j=%d grid=%d lds=%zu item_tiles=%d%s "
                "items=%d "
                "segments=%d "
                "graph=%s",
                mfma_name(c.ks), c.nt, c.ks, c.nbg, c.stage_j ? 1 : 0, run_grid(b, last), c.lds_bytes, b->last_item_tiles,
                b->item_tiles ? "" : " (balanced)",
                b->last_nitems, b->last_segments, b->blk[0].exec || b->blk[1].exec ? "captured" : "none");
  return BF_OK;
}

int bf_plan_autotune(bf_plan *p, uint8_t *J, const uint8_t *E, void *stream, int iters, char *report,
                     size_t report_len) {
  if (!p || !J || !E) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!p->use_mfma) return fail(BF_ERR_UNSUPPORTED, "the plan does not use the MFMA kernel");
  if (iters <= 0) iters = 10;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const bf_shape s = p->shape;
  p->auto_config = false; // the tuned configuration is used as it is, for every launch
  std::string rep;
  TuneChoice best;
  float best_ms = 1e30f;
  int min_ntp = 1 << 30;
  for (int nt = 1; nt <= kMaxBeamTiles; ++nt) {
    MfmaConfig t;
    if (choose_tiling(s.ndishes, s.nbeams, nt, &t)) min_ntp = std::min(min_ntp, t.ntp);
  }
  for (int nt = 1; nt <= kMaxBeamTiles; ++nt) {
    MfmaConfig t;
    if (!choose_tiling(s.ndishes, s.nbeams, nt, &t)) continue;
    if (3 * t.ntp > 4 * min_ntp) continue; // pads a third more beam tiles than the best tiling
    // LDS layout: with the J staging area where it fits, and without it
    MfmaConfig probe = t;
    choose_lds_layout(s.j_layout, true, &probe);
    const int layouts = probe.stage_j ? 2 : 1;
    for (int layout = 0; layout < layouts; ++layout)
      for (double scale : {0.5, 1.0, 2.0}) {
        const TuneChoice choice{nt, scale, layout == 0};
        MfmaConfig c;
        if (!choose_mfma_config(s, p->num_cus, &c, &choice)) continue;
        p->cfg = c;
        if (int rc = alloc_weight_tables(p)) return rc;
        if (int rc = prepare_weights(p, st)) return rc;
        float ms = 0;
        if (int rc = time_launches(st, iters, &ms, [&] { return launch_form_beams(p, J, E, st, false); })) return rc;
        char line[160];
        std::snprintf(line, sizeof line, "nt=%d nbg=%d stage_j=%d scale=%.1f grid=%d lds=%zu: %.1f us\n", c.nt, c.nbg,
                      c.stage_j ? 1 : 0, scale, c.grid, c.lds_bytes, ms * 1e3f);
        rep += line;
        if (ms < best_ms) {
          best_ms = ms;
          best = choice;
        }
      }
  }
  MfmaConfig c;
  if (!choose_mfma_config(s, p->num_cus, &c, &best)) return fail(BF_ERR_UNSUPPORTED, "no candidate configuration");
  p->cfg = c;
  p->wgs_per_cu = wgs_per_cu(c);
  if (int rc = alloc_weight_tables(p)) return rc;
  if (int rc = prepare_weights(p, st)) return rc;
  remember_tuning(s, best);
  char line[160];
  std::snprintf(line, sizeof line, "chosen: nt=%d stage_j=%d scale=%.1f (%.1f us)\n", best.nt, c.stage_j ? 1 : 0,
                best.wg_scale, best_ms * 1e3f);
  rep += line;
  if (report && report_len) std::snprintf(report, report_len, "%s", rep.c_str());
  return BF_OK;
}

int bf_plan_configure(bf_plan *p, int beam_tiles, int stage_j, double wg_scale) {
  if (!p) return fail(BF_ERR_INVALID_ARGUMENT, "NULL plan");
  if (!p->use_mfma) return fail(BF_ERR_UNSUPPORTED, "the plan does not use the MFMA kernel");
  if (beam_tiles < 0 || beam_tiles > kMaxBeamTiles || (stage_j != 0 && stage_j != 1) || !(wg_scale > 0.0) ||
      wg_scale > 64.0)
    return fail(BF_ERR_INVALID_ARGUMENT, "beam_tiles in 0..3, stage_j 0 or 1, wg_scale in (0, 64]");
  const TuneChoice choice{beam_tiles, wg_scale, stage_j != 0};
  MfmaConfig c;
  if (!choose_mfma_config(p->shape, p->num_cus, &c, &choice))
    return fail(BF_ERR_UNSUPPORTED, "no MFMA configuration with these beam tiles for the shape");
  p->auto_config = false; // launches use exactly this configuration
  p->cfg = c;
  p->wgs_per_cu = wgs_per_cu(c);
  if (int rc = alloc_weight_tables(p)) return rc;
  return prepare_weights(p, nullptr);
}

int bf_batch_autotune(bf_batch *b, const bf_chunk *chunks, int nchunks, void *stream, int iters, char *report,
                      size_t report_len) {
  if (!b) return fail(BF_ERR_INVALID_ARGUMENT, "NULL batch");
  if (iters <= 0) iters = 5;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  std::string rep;
  int best_tiles = 0;
  float best_ms = 1e30f;
  const int saved = b->item_tiles;
  constexpr int kCand = 6;
  const int cand[kCand] = {0, 16, 32, 64, 128, 256};
  float sum[kCand] = {};
  int used_tiles[kCand] = {};
  // Pass 0 brings the board to the clock it holds under this load (the
  // first launches after an idle spell run faster) and is not counted;
  // passes 1 and 2 are, with the candidates interleaved so that drifts of
  // clock and temperature reach all of them alike.
  for (int pass = 0; pass < 3; ++pass)
    for (int k = 0; k < kCand; ++k) {
      b->item_tiles = cand[k];
      float ms = 0;
      if (int rc = time_launches(st, iters, &ms, [&] { return bf_batch_run(b, chunks, nchunks, stream); })) {
        b->item_tiles = saved;
        return rc;
      }
      if (pass > 0) sum[k] += ms / 2;
      used_tiles[k] = b->last_item_tiles;
    }
  for (int k = 0; k < kCand; ++k) {
    char line[128];
    std::snprintf(line, sizeof line, "item_tiles=%d%s: %.1f us\n", used_tiles[k], cand[k] ? "" : " (auto)",
                  sum[k] * 1e3f);
    rep += line;
    if (sum[k] < best_ms) {
      best_ms = sum[k];
      best_tiles = cand[k];
    }
  }
  b->item_tiles = best_tiles;
  char line[96];
  std::snprintf(line, sizeof line, "chosen: item_tiles=%d%s\n", best_tiles, best_tiles ? "" : " (auto)");
  rep += line;
  if (report && report_len) std::snprintf(report, report_len, "%s", rep.c_str());
  return BF_OK;
}

} // extern "C"

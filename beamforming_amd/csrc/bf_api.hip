// bf_api.hip -- the C ABI of libbeamform_amd.so (declared in include/beamform.h).
//
// A plan holds everything that depends only on the shape and the weights:
// the weight operands expanded on the device, the per-beam offsets, the
// padded gains and the kernel configuration. form_beams calls are then a
// single launch on the caller's stream.
#include "beamform.h"
#include "bf_plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <array>
#include <map>
#include <mutex>
#include <vector>

namespace bf {

thread_local std::string g_last_error;

int fail(int status, const std::string &msg) {
  g_last_error = msg;
  return status;
}

int hip_fail(hipError_t e, const char *what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? BF_ERR_OUT_OF_MEMORY : BF_ERR_DEVICE;
}

#define BF_HIP(call)                                                                                                   \
  do {                                                                                                                 \
    const hipError_t e_ = (call);                                                                                      \
    if (e_ != hipSuccess) return hip_fail(e_, #call);                                                                  \
  } while (0)

int check_shape(const bf_shape *s) {
  if (!s) return fail(BF_ERR_INVALID_ARGUMENT, "shape is NULL");
  if (s->ntimes <= 0 || s->nfrequencies <= 0 || s->ndishes <= 0 || s->npolarizations <= 0 || s->nbeams <= 0)
    return fail(BF_ERR_INVALID_ARGUMENT, "all shape extents must be positive");
  if (s->j_layout != BF_J_BFPT && s->j_layout != BF_J_BFTP) return fail(BF_ERR_INVALID_ARGUMENT, "bad j_layout");
  if (s->reserved != 0) return fail(BF_ERR_INVALID_ARGUMENT, "reserved field must be 0");
  if (s->nfrequencies > (1 << 20) || s->nbeams > (1 << 20) || s->ndishes > (1 << 20) || s->npolarizations > 64)
    return fail(BF_ERR_UNSUPPORTED, "shape extent too large");
  // int32 accumulation is exact for |raw| <= 128 D < 2^31 and the int -> float
  // conversion of raw is exact below 2^24 (as in the reference).
  if (s->ndishes > 65536) return fail(BF_ERR_UNSUPPORTED, "ndishes > 65536");
  // byte counts of E and J (and time tile indices) stay far inside int64
  const double nE = (double)s->ntimes * (double)s->nfrequencies * (double)s->ndishes * (double)s->npolarizations;
  const double nJ = (double)s->nbeams * (double)s->nfrequencies * (double)s->npolarizations * (double)s->ntimes;
  if (s->ntimes > kMaxTimes || nE > 0x1p62 || nJ > 0x1p62) return fail(BF_ERR_UNSUPPORTED, "E or J too large");
  return BF_OK;
}

// Largest |G| for which |G * raw| < 2^31 holds for every reachable raw
// (|raw| <= 128 D), with a factor-2 margin for the float rounding.
static bool gains_fast_safe(const float *G, int64_t n, int64_t D) {
  const double limit = std::ldexp(1.0, 30) / (128.0 * (double)D);
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(G[i]) || std::fabs((double)G[i]) >= limit) return false;
  return true;
}

static MfmaKernel lookup_mfma(int nt, int ks, bool dt, int jl) {
  switch (ks) {
  case 2: return mfma_kernel_ks2(nt, dt, jl);
  case 4: return mfma_kernel_ks4(nt, dt, jl);
  case 8: return mfma_kernel_ks8(nt, dt, jl);
  case 16: return mfma_kernel_ks16(nt, dt, jl);
  case 32: return mfma_kernel_ks32(nt, dt, jl);
  case 64: return mfma_kernel_ks64(nt, dt, jl);
  default: return nullptr;
  }
}

BatchKernel lookup_batch(int nt, int ks, bool dt, int jl) {
  switch (ks) {
  case 2: return batch_kernel_ks2(nt, dt, jl);
  case 4: return batch_kernel_ks4(nt, dt, jl);
  case 8: return batch_kernel_ks8(nt, dt, jl);
  case 16: return batch_kernel_ks16(nt, dt, jl);
  case 32: return batch_kernel_ks32(nt, dt, jl);
  case 64: return batch_kernel_ks64(nt, dt, jl);
  default: return nullptr;
  }
}

// Tuned choices per shape, filled by bf_plan_autotune and consulted by later
// plans of the same shape in this process.
static std::mutex g_tune_mutex;
static std::map<std::array<int64_t, 5>, TuneChoice> g_tune_cache;

static std::array<int64_t, 5> tune_key(const bf_shape &s) {
  return {s.ntimes, s.nfrequencies, s.ndishes, s.nbeams, s.j_layout};
}

void remember_tuning(const bf_shape &s, const TuneChoice &c) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  g_tune_cache[tune_key(s)] = c;
}

bool recall_tuning(const bf_shape &s, TuneChoice *c) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  auto it = g_tune_cache.find(tune_key(s));
  if (it == g_tune_cache.end()) return false;
  *c = it->second;
  return true;
}

// Beam tiling (KS, NT, beam groups) for D and B. forced_nt > 0 selects NT.
bool choose_tiling(int64_t ndishes, int64_t nbeams, int forced_nt, MfmaConfig *cfg) {
  const int kneed = (int)((ndishes + kDishesPerStep - 1) / kDishesPerStep);
  int ks = 2; // K steps are consumed in pairs
  while (ks < kneed) ks *= 2;
  if (ks > 64) return false;
  const int nt_tot = (int)((nbeams + kBeamTile - 1) / kBeamTile);
  // LDS: ks * nt KiB of weights <= 128 KiB; registers cap nt at 3 (at 4 the
  // accumulators alone take 128 of a wave's 256 registers, and both MFMA
  // forms spill; the int8 form ran slower at 4 than at 2 with twice the beam
  // groups, 242 against 217 us for T=8192 F=16 D=512 B=512)
  const int max_nt = ks >= 64 ? 2 : 3;
  int best_nt = 0, best_nbg = 0;
  if (forced_nt > 0) {
    if (forced_nt > max_nt) return false;
    best_nt = forced_nt;
    best_nbg = (nt_tot + forced_nt - 1) / forced_nt;
  } else {
    // Fewest padded beam tiles, then fewest beam groups (E passes).
    for (int nt = 1; nt <= max_nt; ++nt) {
      const int nbg = (nt_tot + nt - 1) / nt;
      const int padded = nbg * nt;
      if (best_nt == 0 || padded < best_nt * best_nbg || (padded == best_nt * best_nbg && nbg < best_nbg)) {
        best_nt = nt;
        best_nbg = nbg;
      }
    }
  }
  cfg->ks = ks;
  cfg->nt = best_nt;
  cfg->nbg = best_nbg;
  cfg->ntp = best_nt * best_nbg;
  cfg->dish_tail = (int64_t)ks * kDishesPerStep != ndishes;
  choose_lds_layout(BF_J_BFPT, true, cfg); // callers set the plan's layout next
  return cfg->kernel != nullptr && cfg->batch_kernel != nullptr;
}

// LDS layout: the J staging area is used (with either J layout) when it fits
// beside the weights and when wanted; without it the workgroup needs less
// LDS, which can let more workgroups share a CU.
void choose_lds_layout(int j_layout, bool want_staging, MfmaConfig *cfg) {
  // both layouts stage whole rows: (beam, polarization) x 128 B or beam x 256 B
  cfg->stage_j = want_staging && mfma_stages_j(cfg->nt, cfg->ks);
  cfg->lds_bytes = mfma_lds_unstaged(cfg->nt, cfg->ks) + (cfg->stage_j ? mfma_lds_staging(cfg->nt) : 0);
  // the kernels of the configuration, with the J layout's epilogue
  cfg->kernel = lookup_mfma(cfg->nt, cfg->ks, cfg->dish_tail, j_layout);
  cfg->batch_kernel = lookup_batch(cfg->nt, cfg->ks, cfg->dish_tail, j_layout);
}

// Grid for T time samples and F frequencies: about target_wgs workgroups, each
// a multiple of 8 tiles (one per wave).
bool choose_grid(int64_t ntimes, int64_t nfrequencies, int64_t target_wgs, MfmaConfig *cfg) {
  const int64_t ntiles = (ntimes + kTileT - 1) / kTileT;
  const int64_t cols = nfrequencies * cfg->nbg;
  const int64_t target = std::max<int64_t>(target_wgs, 1);
  const int64_t chunks0 = std::max<int64_t>(1, (target + cols - 1) / cols);
  int64_t tpc = (ntiles + chunks0 - 1) / chunks0;
  tpc = (tpc + kWaves - 1) / kWaves * kWaves;
  const int64_t nchunks = (ntiles + tpc - 1) / tpc;
  if (cols * nchunks > (1ll << 30) || tpc > (1ll << 30)) return false;
  cfg->tiles_per_chunk = (int)tpc;
  cfg->nchunks = (int)nchunks;
  cfg->nwork = (int)(cols * nchunks);
  cfg->per_xcd = (cfg->nwork + 7) / 8;
  cfg->grid = 8 * cfg->per_xcd;
  return true;
}

// Resident workgroups per CU of an MFMA kernel configuration.
int wgs_per_cu(const MfmaConfig &cfg) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(cfg.kernel), kThreads,
                                                   cfg.lds_bytes) != hipSuccess ||
      n < 1)
    n = 1;
  return n;
}

// Chooses the MFMA configuration for a shape (tuned choice if one is known).
// Returns false if the shape is outside the MFMA kernel's range.
bool choose_mfma_config(const bf_shape &s, int num_cus, MfmaConfig *cfg, const TuneChoice *force) {
  if (s.npolarizations != 2 || s.ndishes % 8 != 0 || s.ndishes > 1024) return false;
  TuneChoice tc;
  const bool tuned = force ? (tc = *force, true) : recall_tuning(s, &tc);
  if (!choose_tiling(s.ndishes, s.nbeams, tuned ? tc.nt : 0, cfg)) return false;
  choose_lds_layout(s.j_layout, tuned ? tc.stage_j : true, cfg);
  // the instantiation's largest layout, so plans of either layout can share it
  if (hipFuncSetAttribute(reinterpret_cast<const void *>(cfg->kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)mfma_lds_bytes(cfg->nt, cfg->ks)) != hipSuccess)
    return false;
  const double scale = tuned ? tc.wg_scale : 1.0;
  const int64_t target = (int64_t)std::llround(num_cus * wgs_per_cu(*cfg) * scale);
  return choose_grid(s.ntimes, s.nfrequencies, target, cfg);
}

static int num_cus_of_current_device(int *dev, int *ncu) {
  int d = 0;
  BF_HIP(hipGetDevice(&d));
  int n = 0;
  BF_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, d));
  *dev = d;
  *ncu = n;
  return BF_OK;
}

// (Re)builds the device-side weight data of a plan from device A and G.
int prepare_weights(bf_plan *p, hipStream_t st) {
  const bf_shape &s = p->shape;
  const int F = (int)s.nfrequencies, B = (int)s.nbeams, D = (int)s.ndishes;
  if (p->use_mfma) {
    const MfmaConfig &c = p->cfg;
    const int64_t nw = (int64_t)F * c.ks * c.ntp * 64;
    const int blocks = (int)std::min<int64_t>((nw + 255) / 256, 65536);
    hipLaunchKernelGGL(prep_weights_kernel, dim3(blocks), dim3(256), 0, st, p->A_dev, p->W, F, B, D, c.ks, c.ntp);
    BF_HIP(hipGetLastError());
    const int BP = c.ntp * kBeamTile;
    const int64_t nc = (int64_t)F * BP;
    const int blocks2 = (int)std::min<int64_t>((nc + 255) / 256, 65536);
    hipLaunchKernelGGL(prep_corr_kernel, dim3(blocks2), dim3(256), 0, st, p->A_dev, p->G_dev, p->corr, p->Gp, F, B, D,
                       BP);
    BF_HIP(hipGetLastError());
  }
  BF_HIP(hipStreamSynchronize(st));
  return BF_OK;
}

void destroy_host_path(HostPath *hp) {
  if (!hp) return;
  for (hipStream_t st : hp->st)
    if (st) (void)hipStreamSynchronize(st);
  if (hp->slabs) (void)hipFree(hp->slabs);
  for (int k = 0; k < 2; ++k)
    for (hipEvent_t ev : {hp->ev_in[k], hp->ev_run[k], hp->ev_out[k]})
      if (ev) (void)hipEventDestroy(ev);
  for (hipStream_t st : hp->st)
    if (st) (void)hipStreamDestroy(st);
  delete hp;
}

// The plan's host-path resources, made on first use: about 8 slabs of whole
// 16-sample tiles (at least 256 samples) per call, two slab buffers in one
// allocation (E and J of a slab each, 256-byte aligned).
static std::mutex g_host_path_mutex;
int host_path(bf_plan *p, HostPath **out) {
  std::lock_guard<std::mutex> lock(g_host_path_mutex);
  if (p->host) {
    *out = p->host;
    return BF_OK;
  }
  const bf_shape &s = p->shape;
  const int64_t T = s.ntimes;
  HostPath *hp = new HostPath();
  hp->slab = T <= 256 ? T : std::max<int64_t>(256, ((T + 7) / 8 + 15) / 16 * 16);
  hp->nslab = (T + hp->slab - 1) / hp->slab;
  hp->nbuf = hp->nslab > 1 ? 2 : 1;
  const size_t e_row = (size_t)(s.nfrequencies * s.ndishes * s.npolarizations);
  const size_t j_row = (size_t)(s.nbeams * s.nfrequencies * s.npolarizations);
  const size_t e_bytes = ((size_t)hp->slab * e_row + 255) / 256 * 256, j_bytes = ((size_t)hp->slab * j_row + 255) / 256 * 256;
  hipError_t e = hipSuccess;
  int prev = -1;
  const bool switch_dev = hipGetDevice(&prev) == hipSuccess && prev != p->device;
  if (switch_dev) e = hipSetDevice(p->device);
  for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipStreamCreateWithFlags(&hp->st[k], hipStreamNonBlocking);
  for (int k = 0; k < 2 && e == hipSuccess; ++k) {
    e = hipEventCreateWithFlags(&hp->ev_in[k], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&hp->ev_run[k], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&hp->ev_out[k], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipMalloc(&hp->slabs, hp->nbuf * (e_bytes + j_bytes));
  if (switch_dev) (void)hipSetDevice(prev);
  if (e != hipSuccess) {
    destroy_host_path(hp);
    return hip_fail(e, "host path set-up");
  }
  for (int k = 0; k < hp->nbuf; ++k) {
    hp->dE[k] = hp->slabs + k * (e_bytes + j_bytes);
    hp->dJ[k] = hp->dE[k] + e_bytes;
  }
  p->host = hp;
  *out = hp;
  return BF_OK;
}

static void free_plan(bf_plan *p) {
  if (!p) return;
  if (p->W) (void)hipFree(p->W);
  if (p->corr) (void)hipFree(p->corr);
  if (p->Gp) (void)hipFree(p->Gp);
  if (p->A_dev) (void)hipFree(p->A_dev);
  if (p->G_dev) (void)hipFree(p->G_dev);
  destroy_host_path(p->host);
  delete p;
}

// Creates a plan. A and G are host pointers (host_inputs) or device pointers.
static int plan_create(const bf_shape *shape, const uint8_t *A, const float *G, bool host_inputs, hipStream_t st,
                       bf_plan **out) {
  if (!out) return fail(BF_ERR_INVALID_ARGUMENT, "plan output pointer is NULL");
  *out = nullptr;
  if (int rc = check_shape(shape)) return rc;
  if (!A || !G) return fail(BF_ERR_INVALID_ARGUMENT, "A or G is NULL");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(BF_ERR_NO_DEVICE, "no HIP device visible");

  bf_plan *p = new bf_plan();
  p->shape = *shape;
  int rc = BF_OK;
  auto bail = [&](int code) {
    free_plan(p);
    return code;
  };
  if ((rc = num_cus_of_current_device(&p->device, &p->num_cus))) return bail(rc);
  const bf_shape &s = p->shape;
  const int64_t nA = s.nfrequencies * s.nbeams * s.ndishes;
  const int64_t nG = s.nfrequencies * s.nbeams;

  std::vector<float> Gh;
  const float *Gcheck = G;
  if (!host_inputs) {
    Gh.resize(nG);
    hipError_t e = hipMemcpyAsync(Gh.data(), G, nG * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return bail(hip_fail(e, "copy G for range check"));
    Gcheck = Gh.data();
  }
  p->fast_gains = gains_fast_safe(Gcheck, nG, s.ndishes);
  {
    TuneChoice known;
    p->auto_config = !recall_tuning(s, &known); // a tuned choice is kept as it is
  }
  p->use_mfma = p->fast_gains && choose_mfma_config(s, p->num_cus, &p->cfg, nullptr);
  if (p->use_mfma) p->wgs_per_cu = wgs_per_cu(p->cfg);

  hipError_t e;
  if ((e = hipMalloc(&p->A_dev, nA)) != hipSuccess) return bail(hip_fail(e, "hipMalloc A"));
  if ((e = hipMalloc(&p->G_dev, nG * sizeof(float))) != hipSuccess) return bail(hip_fail(e, "hipMalloc G"));
  const hipMemcpyKind kind = host_inputs ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  if ((e = hipMemcpyAsync(p->A_dev, A, nA, kind, st)) != hipSuccess) return bail(hip_fail(e, "copy A"));
  if ((e = hipMemcpyAsync(p->G_dev, G, nG * sizeof(float), kind, st)) != hipSuccess) return bail(hip_fail(e, "copy G"));
  if (p->use_mfma && (rc = alloc_weight_tables(p))) return bail(rc);
  if ((rc = prepare_weights(p, st))) return bail(rc);
  *out = p;
  return BF_OK;
}

// Allocates the weight-operand tables of p->cfg (freeing older ones).
int alloc_weight_tables(bf_plan *p) {
  const bf_shape &s = p->shape;
  const MfmaConfig &c = p->cfg;
  if (p->W) (void)hipFree(p->W);
  if (p->corr) (void)hipFree(p->corr);
  if (p->Gp) (void)hipFree(p->Gp);
  p->W = nullptr;
  p->corr = nullptr;
  p->Gp = nullptr;
  ++p->layout_version; // batches re-derive their launch from the plan
  // the one-beam-tile configuration over the same tables (NTP unchanged)
  p->has_narrow = false;
  if (p->auto_config && c.nt > 1) {
    MfmaConfig n;
    if (choose_tiling(s.ndishes, s.nbeams, 1, &n) && n.ks == c.ks && n.dish_tail == c.dish_tail) {
      n.nbg = n.ntp = c.ntp;
      choose_lds_layout(s.j_layout, true, &n);
      if (hipFuncSetAttribute(reinterpret_cast<const void *>(n.kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)mfma_lds_bytes(n.nt, n.ks)) == hipSuccess) {
        p->cfg_narrow = n;
        p->narrow_wgs_per_cu = wgs_per_cu(n);
        p->has_narrow = true;
      }
    }
  }
  const size_t wbytes = (size_t)s.nfrequencies * c.ks * c.ntp * 64 * sizeof(v4i);
  const size_t BP = (size_t)c.ntp * kBeamTile;
  BF_HIP(hipMalloc(&p->W, wbytes));
  BF_HIP(hipMalloc(&p->corr, s.nfrequencies * BP * 2 * sizeof(int)));
  BF_HIP(hipMalloc(&p->Gp, s.nfrequencies * BP * sizeof(float)));
  return BF_OK;
}

// Launches the MFMA kernel of cfg over E/J with T times and F frequencies
// whose weights are rows wrow0 .. wrow0 + F - 1 of the plan's tables.
int launch_mfma(const bf_plan *p, const MfmaConfig &c, uint8_t *J, const uint8_t *E, int64_t T, int F, int wrow0,
                hipStream_t st) {
  const bf_shape &s = p->shape;
  MfmaParams a;
  a.E = E;
  a.J = J;
  a.W = p->W;
  a.corr = p->corr;
  a.G = p->Gp;
  a.T = T;
  a.F = F;
  a.D = (int32_t)s.ndishes;
  a.B = (int32_t)s.nbeams;
  a.NTP = c.ntp;
  a.nbg = c.nbg;
  a.nchunks = c.nchunks;
  a.tiles_per_chunk = c.tiles_per_chunk;
  a.nwork = c.nwork;
  a.per_xcd = c.per_xcd;
  a.jlayout = s.j_layout;
  a.wrow0 = wrow0;
  a.stage_j = c.stage_j ? 1 : 0;
  void *args[] = {&a};
  BF_HIP(hipLaunchKernel(reinterpret_cast<const void *>(c.kernel), dim3(c.grid), dim3(kThreads), args, c.lds_bytes, st));
  return BF_OK;
}

// Serial work of a launch in units of one beam tile x one round of 8 time
// tiles on one CU: rounds per workgroup x beam tiles x workgroups per CU.
static int64_t launch_cost(const MfmaConfig &c, int num_cus) {
  const int64_t per_cu = (c.nwork + num_cus - 1) / num_cus;
  return per_cu * ((c.tiles_per_chunk + kWaves - 1) / kWaves) * c.nt;
}

// The MFMA configuration of a launch over T times and F frequencies: the
// plan's, with its grid for (T, F), or, when that leaves CUs without a
// workgroup, the one-beam-tile configuration if its serial work is clearly
// smaller (each workgroup does 1/NT of the MFMA work, but every beam group
// reads and unpacks E again: a 1024-sample chunk of one channel at D = 512,
// B = 96 takes 6.4 us instead of 11.9 us back to back, while a launch that
// leaves only a few CUs idle keeps the plan's tiling).
bool pick_config(const bf_plan *p, int64_t T, int64_t F, MfmaConfig *c) {
  *c = p->cfg;
  if (!choose_grid(T, F, (int64_t)p->num_cus * p->wgs_per_cu, c)) return false;
  if (p->has_narrow && c->nwork < p->num_cus) {
    MfmaConfig n = p->cfg_narrow;
    if (choose_grid(T, F, (int64_t)p->num_cus * p->narrow_wgs_per_cu, &n) && n.nwork > c->nwork &&
        5 * launch_cost(n, p->num_cus) < 4 * launch_cost(*c, p->num_cus))
      *c = n;
  }
  return true;
}

int launch_form_beams(const bf_plan *p, uint8_t *J, const uint8_t *E, hipStream_t st, bool allow_narrow) {
  const bf_shape &s = p->shape;
  // The MFMA kernel takes E and J at any byte address (unaligned vector
  // accesses); the generic kernel serves only shapes and gains outside its range.
  if (p->use_mfma) {
    if (allow_narrow && p->has_narrow) {
      MfmaConfig c;
      if (!pick_config(p, s.ntimes, s.nfrequencies, &c)) return fail(BF_ERR_UNSUPPORTED, "ntimes too large");
      return launch_mfma(p, c, J, E, s.ntimes, (int)s.nfrequencies, 0, st);
    }
    return launch_mfma(p, p->cfg, J, E, s.ntimes, (int)s.nfrequencies, 0, st);
  } else {
    const int64_t total = s.nbeams * s.nfrequencies * s.npolarizations * s.ntimes;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(form_beams_generic_kernel, dim3(blocks), dim3(256), 0, st, E, J, p->A_dev, p->G_dev, s.ntimes,
                       (int)s.nfrequencies, (int)s.ndishes, (int)s.npolarizations, (int)s.nbeams, s.j_layout);
    BF_HIP(hipGetLastError());
  }
  return BF_OK;
}

// Device Adler-32 of len bytes with `scratch` (kAdlerScratchWords words of
// device memory): per-block partial pairs, their fold into the checksum in
// the scratch's last words, then the 4-byte result copied to *out once the
// stream has run it. Shared by bf_adler32_device (pooled scratch) and
// bf_adler32_device_static (the library's static device scratch).
int adler32_run(const uint8_t *data, size_t len, unsigned long long *scratch, uint32_t *out, hipStream_t st) {
  const int blocks = adler32_blocks(data, (uint64_t)len);
  uint32_t *result = reinterpret_cast<uint32_t *>(scratch + 2 * kAdlerBlocks);
  hipLaunchKernelGGL(adler32_partial_kernel, dim3(blocks), dim3(kAdlerThreads), 0, st, data, (uint64_t)len, scratch);
  BF_HIP(hipGetLastError());
  hipLaunchKernelGGL(adler32_final_kernel, dim3(1), dim3(kAdlerThreads), 0, st, scratch, blocks, (uint64_t)len, result);
  BF_HIP(hipGetLastError());
  uint32_t h = 0;
  hipError_t e = hipMemcpyAsync(&h, result, sizeof h, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "adler32 device reduction");
  *out = h;
  return BF_OK;
}

// Scratch of bf_adler32_device: a pool of device blocks per device, taken for
// the length of one call and given back, so concurrent calls get distinct
// blocks and steady use allocates nothing. Blocks are kept until the process
// ends.
static std::mutex g_adler_mutex;
static std::vector<std::pair<int, unsigned long long *>> g_adler_pool; // (device, block)

int adler_scratch_take(unsigned long long **out) {
  int dev = 0;
  BF_HIP(hipGetDevice(&dev));
  {
    std::lock_guard<std::mutex> lock(g_adler_mutex);
    for (size_t i = 0; i < g_adler_pool.size(); ++i)
      if (g_adler_pool[i].first == dev) {
        *out = g_adler_pool[i].second;
        g_adler_pool.erase(g_adler_pool.begin() + (std::ptrdiff_t)i);
        return BF_OK;
      }
  }
  BF_HIP(hipMalloc(out, kAdlerScratchWords * sizeof(unsigned long long)));
  return BF_OK;
}

void adler_scratch_give(unsigned long long *p) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return; // dropped: leaked rather than reused on the wrong device
  std::lock_guard<std::mutex> lock(g_adler_mutex);
  g_adler_pool.emplace_back(dev, p);
}

} // namespace bf

using namespace bf;

extern "C" {

const char *bf_status_string(int status) {
  switch (status) {
  case BF_OK: return "ok";
  case BF_ERR_INVALID_ARGUMENT: return "invalid argument";
  case BF_ERR_UNSUPPORTED: return "unsupported shape";
  case BF_ERR_DEVICE: return "device error";
  case BF_ERR_OUT_OF_MEMORY: return "out of memory";
  case BF_ERR_OVER_BUDGET: return "over HBM budget";
  case BF_ERR_NOT_READY: return "not ready";
  case BF_ERR_CLOSED: return "closed";
  case BF_ERR_NO_DEVICE: return "no device";
  default: return "unknown status";
  }
}

const char *bf_version(void) { return "beamforming_amd 0.1 (gfx950, int8 MFMA)"; }

const char *bf_last_error(void) { return g_last_error.c_str(); }

int bf_plan_create(const bf_shape *shape, const uint8_t *A, const float *G, bf_plan **plan) {
  return plan_create(shape, A, G, true, nullptr, plan);
}

int bf_plan_create_device(const bf_shape *shape, const uint8_t *A_dev, const float *G_dev, void *stream,
                          bf_plan **plan) {
  return plan_create(shape, A_dev, G_dev, false, reinterpret_cast<hipStream_t>(stream), plan);
}

int bf_plan_update_weights(bf_plan *p, const uint8_t *A, const float *G) {
  if (!p || !A || !G) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  const bf_shape s = p->shape;
  const int64_t nG = s.nfrequencies * s.nbeams;
  const bool fast = gains_fast_safe(G, nG, s.ndishes);
  if (fast != p->fast_gains) {
    // The kernel family changes; rebuild the plan in place.
    bf_plan *q = nullptr;
    if (int rc = plan_create(&s, A, G, true, nullptr, &q)) return rc;
    const uint64_t version = p->layout_version;
    std::swap(*p, *q);
    free_plan(q);
    p->layout_version = version + 1; // new tables and possibly another kernel family
    return BF_OK;
  }
  BF_HIP(hipMemcpy(p->A_dev, A, s.nfrequencies * s.nbeams * s.ndishes, hipMemcpyHostToDevice));
  BF_HIP(hipMemcpy(p->G_dev, G, nG * sizeof(float), hipMemcpyHostToDevice));
  return prepare_weights(p, nullptr);
}

int bf_plan_destroy(bf_plan *p) {
  free_plan(p);
  return BF_OK;
}

int bf_plan_shape(const bf_plan *p, bf_shape *shape) {
  if (!p || !shape) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  *shape = p->shape;
  return BF_OK;
}

int bf_plan_describe(const bf_plan *p, char *buf, size_t len) {
  if (!p || !buf || len == 0) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (p->use_mfma) {
    const MfmaConfig &c = p->cfg;
    std::snprintf(buf, len,
                  "%s nt=%d ks=%d nbg=%d stage_j=%d tile_t=%d waves=%d tiles_per_chunk=%d nchunks=%d "
                  "grid=%d lds=%zu",
                  mfma_name(c.ks), c.nt, c.ks, c.nbg, c.stage_j ? 1 : 0, kTileT, kWaves, c.tiles_per_chunk, c.nchunks, c.grid,
                  c.lds_bytes);
    if (p->has_narrow) {
      const size_t used = std::strlen(buf);
      if (used + 1 < len) std::snprintf(buf + used, len - used, "; launches under one workgroup per CU: nt=1");
    }
  } else {
    std::snprintf(buf, len, "generic (%s)", p->fast_gains ? "shape outside MFMA range" : "gain outside fast range");
  }
  return BF_OK;
}

int bf_form_beams_device(const bf_plan *p, uint8_t *J, const uint8_t *E, void *stream) {
  if (!p || !J || !E) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  return launch_form_beams(p, J, E, reinterpret_cast<hipStream_t>(stream));
}

// Host buffers: the time axis is cut into slabs that go through two device
// slab buffers, with the copies in, the kernels and the copies out on three
// streams, so that E streams in back to back (the PCIe-bound part) while
// earlier slabs are beamformed and their J goes out (full duplex). A slab's
// J is [b][f][p][n] (or [b][f][n][p]); one 2-D copy puts its rows into the
// columns t0 .. t0 + n - 1 of the caller's J. The streams, events and slab
// buffers are made on the first call and kept with the plan.
int bf_form_beams_host(const bf_plan *p, uint8_t *J, const uint8_t *E) {
  if (!p || !J || !E) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  const bf_shape &s = p->shape;
  const int64_t T = s.ntimes;
  const size_t e_row = (size_t)(s.nfrequencies * s.ndishes * s.npolarizations); // E bytes per time sample
  HostPath *hp = nullptr;
  if (int rc = host_path(const_cast<bf_plan *>(p), &hp)) return rc;
  std::lock_guard<std::mutex> lock(hp->mu);
  enum { IN, RUN, OUT };
  hipStream_t *st = hp->st;
  int rc = BF_OK;
  auto check = [&](hipError_t err, const char *what) {
    if (err != hipSuccess && rc == BF_OK) rc = hip_fail(err, what);
    return rc == BF_OK;
  };
  const bool tp = s.j_layout == BF_J_BFTP; // J rows of n samples x P, else of n samples
  const size_t pitch = tp ? (size_t)T * s.npolarizations : (size_t)T;
  const size_t rows = tp ? (size_t)(s.nbeams * s.nfrequencies) : (size_t)(s.nbeams * s.nfrequencies * s.npolarizations);
  const int nbuf = hp->nbuf;
  for (int64_t i = 0; i < hp->nslab && rc == BF_OK; ++i) {
    const int k = (int)(i % nbuf);
    const bool reuse = i >= nbuf; // slab i - 2 used these buffers (an earlier call has finished with them)
    const int64_t t0 = i * hp->slab, n = std::min(hp->slab, T - t0);
    if (reuse && !check(hipStreamWaitEvent(st[IN], hp->ev_run[k], 0), "hipStreamWaitEvent")) break;
    if (!check(hipMemcpyAsync(hp->dE[k], E + (size_t)t0 * e_row, (size_t)n * e_row, hipMemcpyHostToDevice, st[IN]),
               "copy E slab") ||
        !check(hipEventRecord(hp->ev_in[k], st[IN]), "hipEventRecord") ||
        !check(hipStreamWaitEvent(st[RUN], hp->ev_in[k], 0), "hipStreamWaitEvent") ||
        (reuse && !check(hipStreamWaitEvent(st[RUN], hp->ev_out[k], 0), "hipStreamWaitEvent")))
      break;
    if ((rc = bf_form_beams_channels(p, 0, (int)s.nfrequencies, n, hp->dJ[k], hp->dE[k], st[RUN])) != BF_OK) break;
    const size_t width = tp ? (size_t)n * s.npolarizations : (size_t)n;
    const size_t col = tp ? (size_t)t0 * s.npolarizations : (size_t)t0;
    if (!check(hipEventRecord(hp->ev_run[k], st[RUN]), "hipEventRecord") ||
        !check(hipStreamWaitEvent(st[OUT], hp->ev_run[k], 0), "hipStreamWaitEvent") ||
        !check(hipMemcpy2DAsync(J + col, pitch, hp->dJ[k], width, width, rows, hipMemcpyDeviceToHost, st[OUT]),
               "copy J slab") ||
        !check(hipEventRecord(hp->ev_out[k], st[OUT]), "hipEventRecord"))
      break;
  }
  for (int k = 0; k < 3; ++k) {
    const hipError_t e = hipStreamSynchronize(st[k]);
    if (e != hipSuccess && rc == BF_OK) rc = hip_fail(e, "host path");
  }
  return rc;
}

int bf_form_beams(uint8_t *J, const uint8_t *E, const uint8_t *A, const float *G, const bf_shape *shape) {
  bf_plan *p = nullptr;
  if (int rc = bf_plan_create(shape, A, G, &p)) return rc;
  const int rc = bf_form_beams_host(p, J, E);
  bf_plan_destroy(p);
  return rc;
}

uint32_t bf_adler32(const uint8_t *data, size_t len) {
  // Same recurrence as adler32.c:7-20, with the modulo deferred over blocks
  // of 5552 bytes (the largest count for which b cannot overflow 32 bits).
  const uint32_t mod = 65521u;
  uint32_t a = 1, b = 0;
  while (len > 0) {
    const size_t n = len < 5552 ? len : 5552;
    for (size_t i = 0; i < n; ++i) {
      a += data[i];
      b += a;
    }
    a %= mod;
    b %= mod;
    data += n;
    len -= n;
  }
  return (b << 16) | a;
}

int bf_adler32_device(const uint8_t *data, size_t len, uint32_t *out, void *stream) {
  if (!data || !out) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  unsigned long long *scratch = nullptr;
  if (int rc = adler_scratch_take(&scratch)) return rc;
  const int rc = adler32_run(data, len, scratch, out, st);
  adler_scratch_give(scratch);
  return rc;
}

int bf_fill_uniform_device(uint8_t *buf, size_t n, uint64_t seed, uint64_t offset, void *stream) {
  if (!buf && n) return fail(BF_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (n == 0) return BF_OK;
  const uint64_t nvec = n / 16;
  const int blocks = (int)std::max<uint64_t>(1, std::min<uint64_t>((nvec + 255) / 256, 8192));
  hipLaunchKernelGGL(fill_uniform_kernel, dim3(blocks), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), buf,
                     (uint64_t)n, seed, offset);
  BF_HIP(hipGetLastError());
  return BF_OK;
}

} // extern "C"

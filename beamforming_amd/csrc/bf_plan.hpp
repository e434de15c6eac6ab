// bf_plan.hpp -- the plan object behind the C ABI.
#pragma once

#include "beamform.h"
#include "bf_kernels.hpp"

#include <mutex>
#include <string>

namespace bf {

// MFMA kernel configuration for one shape.
struct MfmaConfig {
  int ks = 0;              // K steps (power of two, 16 dishes each)
  bool dish_tail = false;  // ks * 16 != ndishes: masked last steps
  int nt = 0;              // 32-beam tiles per workgroup
  int nbg = 0;             // beam groups
  int ntp = 0;             // nt * nbg
  int tiles_per_chunk = 0; // 16-sample tiles per workgroup
  int nchunks = 0;         // time chunks per (f, beam group)
  int nwork = 0;           // workgroups with work
  int per_xcd = 0;         // work items per XCD
  int grid = 0;            // launched workgroups (8 * per_xcd)
  bool stage_j = false;    // J rows staged in LDS (whole-line stores); else direct 16-byte stores
  size_t lds_bytes = 0;    // dynamic LDS per workgroup
  MfmaKernel kernel = nullptr;
  BatchKernel batch_kernel = nullptr;
};

// A tuned choice for one shape: beam tiles per workgroup and a scale on the
// default number of workgroups.
struct TuneChoice {
  int nt = 0;
  double wg_scale = 1.0;
  bool stage_j = true; // stage J rows in LDS when the layout and the LDS allow it
};

// Resources of the host-buffer entry point (bf_form_beams_host), made on
// its first call and kept with the plan: the copy-in, kernel and copy-out
// streams, the events ordering slab-buffer reuse, and the two slab buffers
// (one device allocation). Calls on one plan are serialised by `mu`.
struct HostPath {
  std::mutex mu;
  hipStream_t st[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_in[2] = {nullptr, nullptr}, ev_run[2] = {nullptr, nullptr}, ev_out[2] = {nullptr, nullptr};
  uint8_t *slabs = nullptr;
  uint8_t *dE[2] = {nullptr, nullptr}, *dJ[2] = {nullptr, nullptr};
  int64_t slab = 0, nslab = 0; // samples per slab, slabs per call
  int nbuf = 0;
};

} // namespace bf

struct bf_plan {
  bf_shape shape{};
  int device = 0;
  int num_cus = 0;
  bool fast_gains = false; // all gains inside the fast requantisation range
  bool use_mfma = false;   // MFMA kernel (else the generic kernel)
  bf::MfmaConfig cfg;
  // Small launches (fewer workgroups than CUs with cfg) take one beam tile
  // per workgroup instead: NT times the workgroups over the same weight
  // tables. Only for the automatic configuration, not a configured or tuned one.
  bool auto_config = false;
  bool has_narrow = false;
  bf::MfmaConfig cfg_narrow;
  int narrow_wgs_per_cu = 1;
  uint64_t layout_version = 0; // bumped whenever cfg or the weight tables' addresses change
  int wgs_per_cu = 1;      // resident MFMA workgroups per CU
  uint8_t *A_dev = nullptr; // [F][B][D] as given
  float *G_dev = nullptr;   // [F][B] as given
  bf::v4i *W = nullptr;     // [F][KS / 2][2 NTP][64] weight operands (32-dish steps, 16-beam groups)
  int *corr = nullptr;      // [F][BP][2]
  float *Gp = nullptr;      // [F][BP]
  bf::HostPath *host = nullptr; // bf_form_beams_host resources (made on first use)
};

namespace bf {
extern thread_local std::string g_last_error;
int fail(int status, const std::string &msg);
int hip_fail(hipError_t e, const char *what);
int check_shape(const bf_shape *s);
bool choose_mfma_config(const bf_shape &s, int num_cus, MfmaConfig *cfg, const TuneChoice *force);
bool choose_tiling(int64_t ndishes, int64_t nbeams, int forced_nt, MfmaConfig *cfg);
void choose_lds_layout(int j_layout, bool want_staging, MfmaConfig *cfg);
bool choose_grid(int64_t ntimes, int64_t nfrequencies, int64_t target_wgs, MfmaConfig *cfg);
int wgs_per_cu(const MfmaConfig &cfg);
void remember_tuning(const bf_shape &s, const TuneChoice &c);
bool recall_tuning(const bf_shape &s, TuneChoice *c);
int alloc_weight_tables(bf_plan *p);
int prepare_weights(bf_plan *p, hipStream_t st);
int launch_mfma(const bf_plan *p, const MfmaConfig &c, uint8_t *J, const uint8_t *E, int64_t T, int F, int wrow0,
                hipStream_t st);
int launch_form_beams(const bf_plan *p, uint8_t *J, const uint8_t *E, hipStream_t st, bool allow_narrow = true);
int adler32_run(const uint8_t *data, size_t len, unsigned long long *scratch, uint32_t *out, hipStream_t st);
bool pick_config(const bf_plan *p, int64_t T, int64_t F, MfmaConfig *c);
} // namespace bf

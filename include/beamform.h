/*
 * beamform.h -- C ABI of the MI355X (gfx950) form-beams library
 * (libbeamform_amd.so).
 *
 * The reference computes, for every frequency f, beam b, polarization p and
 * time t,
 *     J[b,f,p,t] = requant( G[f,b] * sum_d A[f,b,d] * E[t,f,d,p] )
 * on 4-bit complex values (reference cpu.cxx:15-46). Its entry point is
 *     void form_beams(ucomplex4 *J, const ucomplex4 *E, const ucomplex4 *A,
 *                     const float *G)                       (cpu.cxx:15-16)
 * with shapes fixed at compile time (arraysizes.hxx:22-27) and, on the GPU,
 * the same four pointers as device pointers (cuda.cu:35-36, launched at
 * cuda.cu:291-296). This header replaces both: shapes are runtime values, the
 * weights live in a plan, and every function returns a status instead of
 * calling exit() (cuda.cu:255-262, arraysizes.cxx:82-86).
 *
 * Data formats (identical to the reference, byte for byte):
 *   one complex value = one byte, ucomplex4 offset binary: high nibble = real,
 *   low nibble = imaginary, each stored as value + 8 (icomplex4.hxx:22-40);
 *   E[t][f][d][p]  (Elinear, arraysizes.hxx:79-89)
 *   A[f][b][d]     (Alinear, arraysizes.hxx:114-123)
 *   G[f][b] float  (Glinear, arraysizes.hxx:135-142)
 *   J[b][f][p][t]  (Jlinear, arraysizes.hxx:91-101; BF_J_BFPT, the layout of
 *                   every reference checksum) or
 *   J[b][f][t][p]  (J2linear, arraysizes.hxx:103-112; BF_J_BFTP)
 *
 * Streams are hipStream_t values passed as void*; NULL is the null stream.
 * All device pointers are plain hipMalloc'd (or torch) device memory. E and J
 * may start at any byte address: the kernels use unaligned vector accesses.
 * Every E and J buffer must hold the bytes its shape gives (a chunk or a
 * request: exactly its own ntimes samples); the library cannot check sizes.
 */
#ifndef BEAMFORM_H
#define BEAMFORM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
enum {
  BF_OK = 0,
  BF_ERR_INVALID_ARGUMENT = 1, /* null pointer, negative size, bad layout */
  BF_ERR_UNSUPPORTED = 2,      /* shape outside what the library handles */
  BF_ERR_DEVICE = 3,           /* a HIP runtime call failed */
  BF_ERR_OUT_OF_MEMORY = 4,    /* device or host allocation failed */
  BF_ERR_OVER_BUDGET = 5,      /* serving: request larger than the HBM budget */
  BF_ERR_NOT_READY = 6,        /* serving: no finished request to return */
  BF_ERR_CLOSED = 7,           /* serving: server shut down */
  BF_ERR_NO_DEVICE = 8         /* no HIP device visible */
};

/* J layouts. */
enum { BF_J_BFPT = 0, BF_J_BFTP = 1 };

typedef struct bf_shape {
  int64_t ntimes;         /* T: time samples per chunk (arraysizes.hxx:22) */
  int64_t nfrequencies;   /* F: frequency channels (arraysizes.hxx:23) */
  int64_t ndishes;        /* D: dishes (arraysizes.hxx:24) */
  int64_t npolarizations; /* P: must be 2 (arraysizes.hxx:25) */
  int64_t nbeams;         /* B: beams (arraysizes.hxx:26) */
  int32_t j_layout;       /* BF_J_BFPT (default) or BF_J_BFTP */
  int32_t reserved;       /* must be 0 */
} bf_shape;

typedef struct bf_plan bf_plan;

/* Human-readable text for a status code. */
const char *bf_status_string(int status);

/* Library version string. */
const char *bf_version(void);

/* Text of the last error raised on the calling thread ("" if none). */
const char *bf_last_error(void);

/* Creates a plan for `shape` with weights A (F*B*D bytes) and gains G (F*B
 * floats) in HOST memory. The weights are expanded once on the device into
 * the operand layout the kernels read and stay resident with the plan; the
 * kernel configuration for the shape is chosen here. Replaces the per-call
 * cudaMemcpy of A and G in cuda.cu:274-284. */
int bf_plan_create(const bf_shape *shape, const uint8_t *A, const float *G, bf_plan **plan);

/* Same, with A and G already in device memory (stream-ordered on `stream`;
 * the call returns after the preparation has finished). */
int bf_plan_create_device(const bf_shape *shape, const uint8_t *A_dev, const float *G_dev, void *stream,
                          bf_plan **plan);

/* Replaces a plan's weights and gains (host pointers), keeping its shape.
 * Work already enqueued with the plan must have finished. */
int bf_plan_update_weights(bf_plan *plan, const uint8_t *A, const float *G);

/* Frees a plan. NULL is accepted. */
int bf_plan_destroy(bf_plan *plan);

/* Copies the plan's shape out. */
int bf_plan_shape(const bf_plan *plan, bf_shape *shape);

/* Writes a one-line description of the kernel configuration chosen for the
 * plan (kernel family, tile sizes, grid) into buf. */
int bf_plan_describe(const bf_plan *plan, char *buf, size_t len);

/* form_beams on device buffers, asynchronous and ordered on `stream`:
 * J (B*F*P*T bytes) <- E (T*F*D*P bytes). The GPU form of cpu.cxx:15-46 /
 * cuda.cu:35-253; J is fully overwritten. A plan is immutable while in use,
 * so calls on different streams may run concurrently (distinct J buffers).
 * Shapes with P = 2, D a multiple of 8 up to 1024 and gains G with
 * |G| * 128 D < 2^30 run the MFMA kernel (fp6 operands that hold the int4
 * values exactly, for 33 to 512 dishes; int8 otherwise) at any alignment of
 * E and J;
 * other shapes and gains run a one-thread-per-output kernel with the same
 * results (bf_plan_describe names the kernel of a plan). */
int bf_form_beams_device(const bf_plan *plan, uint8_t *J, const uint8_t *E, void *stream);

/* form_beams on host buffers: copies E in, runs, copies J out, synchronously
 * (the flow of cuda.cu:274-314). The time axis goes through the device in
 * slabs, the copies in, the kernels and the copies out on three streams, so
 * that they overlap (best with pinned host buffers); device memory use is
 * two slabs of about T / 8 samples each, not whole E and J. The streams,
 * events and slab buffers are made on a plan's first call and kept with the
 * plan; calls on one plan from several threads run one after the other. */
int bf_form_beams_host(const bf_plan *plan, uint8_t *J, const uint8_t *E);

/* One-shot drop-in for the reference entry point with an explicit shape, all
 * pointers in host memory: plan, run, destroy (cpu.cxx:15-16). */
int bf_form_beams(uint8_t *J, const uint8_t *E, const uint8_t *A, const float *G, const bf_shape *shape);

/* Adler-32 (adler32.c:7-20) of `len` bytes of DEVICE memory, computed on the
 * device; the result is written to *out (host) after the stream has run it.
 * Its 32 KiB of device scratch comes from a per-device pool kept by the
 * library (allocated on first use, reused by later calls; concurrent calls
 * get distinct scratch). */
int bf_adler32_device(const uint8_t *data, size_t len, uint32_t *out, void *stream);

/* Same as bf_adler32_device, allocating nothing: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Its
 * 32 KiB + 16 bytes of scratch is a static device array of the library, one
 * per GPU, held until the process ends. Any len is accepted (the scratch
 * does not grow with it). Calls that overlap on one GPU share that scratch
 * and race; calls one after the other on one stream are safe. */
int bf_adler32_device_noalloc(const uint8_t *data, size_t len, uint32_t *out, void *stream);

/* Adler-32 of host memory (adler32.c:7-20), for the checksum flow of
 * arraysizes.cxx:62-88 without a device round trip. */
uint32_t bf_adler32(const uint8_t *data, size_t len);

/* Fills `n` bytes of device memory with seeded uniform ucomplex4 values:
 * byte i = enc(re, im), re = ((lo32(z) * 15) >> 32) - 7,
 * im = ((hi32(z) * 15) >> 32) - 7, z = SplitMix64 output number (offset + i)
 * for `seed`. Used to build large synthetic inputs in place. */
int bf_fill_uniform_device(uint8_t *buf, size_t n, uint64_t seed, uint64_t offset, void *stream);

/* ---------------------------------------------------------------------------
 * Channel tables, batching and tuning. These have no counterpart in the
 * reference, which runs one compile-time chunk per call (cpu.cxx:48-67,
 * cuda4.cu:1568-1580 launches one kernel per stream); they serve the
 * telescope's uneven per-channel chunks.
 *
 * A "channel-table plan" is an ordinary plan whose frequencies are the
 * channels: A[c][B][D], G[c][B]; its ntimes is not used by the calls below.
 * ------------------------------------------------------------------------- */

/* form_beams for ONE channel of a plan, with its own ntimes: E is
 * [ntimes][1][D][P], J is [B][1][P][ntimes] (or [B][1][ntimes][P]). This is
 * the plain per-chunk launch. Asynchronous on `stream`. */
int bf_form_beams_channel(const bf_plan *plan, int channel, int64_t ntimes, uint8_t *J, const uint8_t *E,
                          void *stream);

/* Same for nfreq consecutive channels channel .. channel + nfreq - 1 in one
 * launch: E is [ntimes][nfreq][D][P], J is [B][nfreq][P][ntimes] (or
 * [B][nfreq][ntimes][P]). */
int bf_form_beams_channels(const bf_plan *plan, int channel, int nfreq, int64_t ntimes, uint8_t *J,
                           const uint8_t *E, void *stream);

typedef struct bf_chunk {
  const uint8_t *E; /* device: [ntimes][nfreq][D][P] (Elinear with F = nfreq), exactly ntimes samples */
  uint8_t *J;       /* device: [B][nfreq][P][ntimes] or [B][nfreq][ntimes][P] per the plan layout,
                       exactly ntimes samples (the kernel addresses rows with this chunk's ntimes) */
  int64_t ntimes;   /* >= 1 */
  int32_t channel;  /* first row of the plan's channel table used by the chunk */
  int32_t nfreq;    /* consecutive channels channel .. channel + nfreq - 1 (0 means 1); <= 65535 */
} bf_chunk;

typedef struct bf_batch bf_batch;

/* Creates a batch context for up to max_chunks chunks of a channel-table
 * plan: two descriptor blocks in pinned host memory (chunk list and work
 * items, written by alternate runs and read by the kernel in place), and a
 * hipGraph of the launch per block, captured on first use of
 * bf_batch_run_graph. The
 * batch follows its plan: after bf_plan_configure, bf_plan_autotune or
 * bf_plan_update_weights the next run re-derives its launch (and captures the
 * graph again); it fails with BF_ERR_UNSUPPORTED while new gains keep the
 * plan off the MFMA kernel. */
int bf_batch_create(const bf_plan *channel_plan, int max_chunks, bf_batch **batch);
int bf_batch_destroy(bf_batch *batch);

/* Fixed work-item size in 16-sample tiles (rounded to a multiple of 8;
 * items taken from the device-side queue longest first). 0 restores the
 * default, balanced items (at most one per workgroup): the
 * chunks' tile ranges laid end to end and cut into one equal share per
 * workgroup of the persistent grid, counted in rounds of 8 tiles (a chunk's
 * last round may be partial), a share crossing chunk boundaries being run as
 * several segments. */
int bf_batch_set_item_tiles(bf_batch *batch, int item_tiles);

/* Runs all chunks in one persistent launch on `stream` (balanced items one
 * per workgroup, fixed-size items through a device-side work queue;
 * asynchronous, the chunk list is read before the call returns).
 * The launch needs no copy: the kernel reads the run's descriptor block from
 * pinned host memory. A run
 * first waits, on the host, until the run two before it on this batch has
 * finished (it reuses that run's block); runs issued on different streams
 * may overlap. A balanced batch too small to give every workgroup an item
 * runs one beam tile per workgroup (plans with the automatic configuration).
 * The plan must use the MFMA path (P = 2, D a multiple of 8 up to 1024, gains
 * in the fast range); otherwise BF_ERR_UNSUPPORTED. */
int bf_batch_run(bf_batch *batch, const bf_chunk *chunks, int nchunks, void *stream);

/* Same work, launched as a replay of a captured hipGraph of the persistent
 * kernel (one graph per descriptor block, captured on first use). */
int bf_batch_run_graph(bf_batch *batch, const bf_chunk *chunks, int nchunks, void *stream);

/* One-line description of the batch configuration. */
int bf_batch_describe(const bf_batch *batch, char *buf, size_t len);

/* Sets the kernel configuration explicitly, as the autotuner does after
 * timing: beam_tiles 32-beam tiles per workgroup (1..3, 0 = automatic),
 * stage_j 1 to stage J rows in LDS where the layout and the LDS allow it,
 * wg_scale times the default number of workgroups. Work already enqueued
 * with the plan must have finished. Synchronous. */
int bf_plan_configure(bf_plan *plan, int beam_tiles, int stage_j, double wg_scale);

/* Times candidate kernel configurations (beam tiles per workgroup, LDS
 * layout: J rows staged in LDS for whole-line stores or stored directly, and
 * the number of workgroups, i.e. the time split) on the caller's device
 * buffers E and J with the plan's shape, keeps the fastest in the plan, and
 * remembers it for later plans of the same shape in this process. `iters`
 * timed launches per candidate (0 = default). A report of the candidates is
 * written to report (may be NULL). Synchronous. */
int bf_plan_autotune(bf_plan *plan, uint8_t *J, const uint8_t *E, void *stream, int iters, char *report,
                     size_t report_len);

/* Times candidate work-item sizes for the given chunk list (balanced items,
 * or fixed sizes of 16 to 256 tiles; one persistent launch each, `iters`
 * times, in three interleaved passes of which the first only warms the
 * board) and keeps the fastest. Synchronous. */
int bf_batch_autotune(bf_batch *batch, const bf_chunk *chunks, int nchunks, void *stream, int iters, char *report,
                      size_t report_len);

/* ---------------------------------------------------------------------------
 * Serving: a streaming request loop with continuous batching. No counterpart
 * in the reference (its drivers run one chunk per process, cpu.cxx:48-67).
 *
 * Requests (one chunk of nfreq consecutive channels of a channel-table plan,
 * with its own ntimes) are submitted from any thread. A loop thread admits
 * them in arrival order while a slot is free and the HBM budget (one device
 * arena allocated at creation) has room for the request's J and, for host-fed
 * E, its staging copy; it launches everything admitted as one persistent
 * batched launch (bf_batch_run; a lone request as bf_form_beams_channels)
 * on one of max_launches streams, so requests
 * arriving while others run are taken into the next launch and overlap the
 * tail of the earlier one. J stays in the arena until bf_server_release.
 * The plan must outlive the server.
 * ------------------------------------------------------------------------- */

typedef struct bf_server bf_server;

typedef struct bf_server_config {
  size_t hbm_budget;    /* bytes of device memory for J outputs and staged E; > 0 */
  int32_t max_inflight; /* requests admitted and not finished at once (0 = 256) */
  int32_t max_batch;    /* requests per launch (0 = 128) */
  int32_t max_launches; /* launches in flight at once, one stream each (0 = 2) */
  int32_t use_graph;    /* 1: each launch of several requests is a replay of a captured hipGraph */
  int32_t poll_us;      /* latency against CPU use. 0 (default): while launches are in flight the
                           loop thread polls their completion without sleeping, yielding its core
                           (one core stays busy while the GPU works), and a submission is launched
                           at once from the submitting thread when a launch slot is free.
                           N in 1..1000000: the loop sleeps up to N us between polls (the OS may wake
                           it tens of microseconds late) and launches everything itself */
} bf_server_config;

typedef struct bf_request {
  const uint8_t *E;  /* [ntimes][nfreq][D][P]: device memory, or host if e_on_host */
  int64_t ntimes;    /* >= 1 */
  int32_t channel;   /* first channel-table row */
  int32_t nfreq;     /* consecutive channels (0 means 1) */
  int32_t e_on_host; /* 1: E is host memory, copied into the arena at launch (must stay valid until the
                        result is returned); 0: device memory, read in place (same lifetime rule) */
  int32_t reserved;  /* must be 0 */
  uint64_t tag;      /* caller's value, returned with the result */
} bf_request;

typedef struct bf_result {
  uint64_t ticket;
  uint64_t tag;
  uint8_t *J;         /* device: [B][nfreq][P][ntimes] (or BFTP), valid until bf_server_release */
  size_t j_bytes;
  int32_t status;     /* BF_OK, or the error of the launch that carried the request */
  int32_t batch_size; /* requests in that launch */
  double t_submit, t_admit, t_launch, t_done; /* seconds on the server clock */
} bf_result;

typedef struct bf_server_counters {
  uint64_t submitted, admitted, completed, released, rejected, failed;
  uint64_t launches, max_batch, waiting, inflight;
  size_t arena_bytes, bytes_in_use, peak_bytes;
  double mean_batch; /* requests per launch */
  double clock;      /* server clock now, seconds */
} bf_server_counters;

/* Creates a server over a channel-table plan (MFMA path required). */
int bf_server_create(const bf_plan *channel_plan, const bf_server_config *config, bf_server **server);

/* Finishes the launches in flight, drops requests not yet launched, frees
 * everything (J buffers of unreleased results included). NULL is accepted. */
int bf_server_destroy(bf_server *server);

/* Queues a request; returns at once. BF_ERR_OVER_BUDGET if the request alone
 * needs more than the whole budget, BF_ERR_INVALID_ARGUMENT for a malformed
 * request, BF_ERR_CLOSED while shutting down. */
int bf_server_submit(bf_server *server, const bf_request *request, uint64_t *ticket);

/* Queues n requests at once, atomically with respect to the loop (they are
 * admitted and launched together as far as budget and slots allow); all or
 * none are queued. tickets (may be NULL) receives n tickets. */
int bf_server_submit_many(bf_server *server, const bf_request *requests, int n, uint64_t *tickets);

/* Returns a finished request (each once), or BF_ERR_NOT_READY. */
int bf_server_poll(bf_server *server, bf_result *result);

/* Waits up to timeout_s seconds (< 0: without limit) for a finished request;
 * BF_ERR_NOT_READY on timeout. */
int bf_server_wait(bf_server *server, double timeout_s, bf_result *result);

/* Copies up to len bytes of a finished request's J to host memory. */
int bf_server_read(bf_server *server, uint64_t ticket, uint8_t *host, size_t len);

/* Returns a finished request's J memory to the budget. */
int bf_server_release(bf_server *server, uint64_t ticket);

int bf_server_stats(bf_server *server, bf_server_counters *counters);

/* Seconds since the server was created (the clock of bf_result). */
double bf_server_clock(const bf_server *server);

/* Replaces the channel table's weights and gains (host pointers, A [C][B][D]
 * and G [C][B] like the plan's) for the requests submitted after this call
 * returns; requests submitted earlier, launched or not, finish with the
 * weights they were submitted under. The new set is expanded on the device
 * while the server keeps launching (double buffering, no pause); a set is
 * freed once none of its requests is pending (at a later update or at
 * shutdown). The caller's plan is never modified. Synchronous; the new gains
 * must keep the MFMA path (else BF_ERR_UNSUPPORTED and nothing changes). The
 * weights/gain update path of the reference's staged variant (cuda4.cu:
 * 1468-1494 rebuilds A2 per run) as a serving operation. */
int bf_server_update_weights(bf_server *server, const uint8_t *A, const float *G);

/* ---------------------------------------------------------------------------
 * E ingest: packets of ntimes time samples for all F channels, in the
 * reference layout E[t][f][d][p] (Elinear, arraysizes.hxx:79-89), are cut
 * into chunks of chunk_ntimes samples for each group of group_channels
 * consecutive channels and written into per-group device ring buffers
 * ([chunk_ntimes][G][D][P] per slot), by strided 2-D copies on the ingest's
 * own stream. A popped chunk is a bf_chunk (nfreq = G, J left NULL) for
 * bf_batch_run or the server. One thread drives an ingest object.
 * ------------------------------------------------------------------------- */

typedef struct bf_ingest bf_ingest;

typedef struct bf_ingest_config {
  int64_t nfrequencies;   /* F: channels in every packet */
  int64_t ndishes;        /* D: multiple of 8 */
  int64_t npolarizations; /* must be 2 */
  int64_t chunk_ntimes;   /* samples per output chunk */
  int32_t group_channels; /* G: channels per output chunk; divides F */
  int32_t ring_slots;     /* chunks buffered per group, >= 2 */
  int32_t reserved;       /* must be 0 */
  int32_t reserved2;      /* must be 0 */
} bf_ingest_config;

int bf_ingest_create(const bf_ingest_config *config, bf_ingest **ingest);
int bf_ingest_destroy(bf_ingest *ingest);

/* Appends a packet of ntimes samples (E[ntimes][F][D][P]; host memory if
 * packet_on_host, best pinned, else device memory). Packets continue each
 * other in time. Returns BF_ERR_NOT_READY, copying nothing, when a chunk it
 * would start has no free ring slot (the consumer has not released one).
 * The packet is read asynchronously, on the ingest's own stream: keep it
 * unchanged until bf_ingest_sync returns (or until a stream made to wait by
 * bf_ingest_pop for a chunk holding the packet's last sample has passed
 * that point). */
int bf_ingest_push(bf_ingest *ingest, const uint8_t *packet, int64_t ntimes, int packet_on_host);

/* Pushes packets[0 .. n-1] (ntimes[i] samples each) in order, as n calls of
 * bf_ingest_push, stopping at the first one refused: its status is returned
 * (BF_ERR_NOT_READY when a ring slot is missing) and *pushed (may be NULL)
 * counts the packets taken. For receivers that get several packets at once. */
int bf_ingest_push_many(bf_ingest *ingest, const uint8_t *const *packets, const int64_t *ntimes, int n,
                        int packet_on_host, int *pushed);

/* Waits until every packet pushed so far has been read; the caller may then
 * reuse the packet buffers. */
int bf_ingest_sync(bf_ingest *ingest);

/* Next complete chunk, in completion order (for every group in turn), or
 * BF_ERR_NOT_READY. `stream` is made to wait for the chunk's copies. */
int bf_ingest_pop(bf_ingest *ingest, void *stream, bf_chunk *chunk, uint64_t *slot_id);

/* Returns a popped chunk's slot to the ring once the work enqueued on
 * `stream` so far (the consumer of the chunk) has finished. */
int bf_ingest_release(bf_ingest *ingest, uint64_t slot_id, void *stream);

/* Blocks until the oldest release (bf_ingest_release) whose consumer work
 * may still be running has finished, and frees its slot: a receiver whose
 * push was refused with BF_ERR_NOT_READY calls it and pushes again, instead
 * of synchronising the whole device, so the copies already queued keep the
 * GPU busy meanwhile. Returns BF_ERR_NOT_READY at once when no release is
 * pending (only the consumer can free a slot then). No counterpart in the
 * reference, which has no ingest. */
int bf_ingest_wait_release(bf_ingest *ingest);

/* Samples in the chunks being filled and complete chunks not yet popped. */
int bf_ingest_pending(const bf_ingest *ingest, int64_t *filled_samples, int *ready_chunks);

#ifdef __cplusplus
}
#endif

#endif /* BEAMFORM_H */

// bf_server.hip -- the streaming request loop (continuous batching).
//
// Requests (one chunk each: E of nfreq consecutive channels, own ntimes) are
// submitted from any thread (with poll_us 0 the submitting thread itself
// admits and launches when a launch slot is free). A loop thread owned by the
// server
//   1. admits waiting requests in arrival order while a slot is free and the
//      HBM arena (the budget, one device allocation made at creation) has
//      room for the request's J and, for host-fed E, its staging copy;
//   2. launches everything admitted so far as ONE persistent batched launch
//      (work queue over all chunks, longest items first; a lone request as a
//      plain channel-range launch) on a free launch stream, up to
//      max_launches launches in flight; a launch issued while
//      another is still draining starts on the CUs that the other's
//      workgroups leave, so new requests overlap the tail of earlier ones;
//   3. polls the launches' events (without sleeping unless poll_us asks for
//      it), stamps the finished requests and hands them to bf_server_poll /
//      bf_server_wait.
// A result's J stays in the arena until bf_server_release; memory returned
// there (or by finished E staging) lets the next waiting requests in.
//
// Weights are double buffered: bf_server_update_weights prepares a new weight
// set (a plan of the same shape and its batches) while requests keep running
// on the current one; requests carry the set that was current when they were
// submitted, a launch never mixes sets, and a set is freed once none of its
// requests is pending (at the next update or at shutdown).
#include "beamform.h"
#include "bf_plan.hpp"
#include "bf_sched.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace bf;

namespace {

struct Request {
  uint64_t ticket = 0, tag = 0;
  const uint8_t *E_src = nullptr;
  bool e_on_host = false;
  int64_t ntimes = 0;
  int channel = 0, nfreq = 1;
  size_t e_bytes = 0, j_bytes = 0;
  size_t j_off = 0, e_off = 0;
  bool has_e_piece = false, has_j_piece = false;
  double t_submit = 0, t_admit = 0, t_launch = 0, t_done = 0;
  int status = BF_OK;
  int batch_size = 0;
  bool done = false;
  uint64_t gen = 0; // weight set
};

struct Launch {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  // busy: the slot is taken (from the moment a thread picks it until its
  // requests are finished). issued: the launch's copies, kernel and `done`
  // event have been enqueued; set under the lock once issue() has returned.
  // Until then hipEventQuery(done) reports an earlier record (or none, which
  // also reads as complete), so the completion scan must not look at it.
  bool busy = false;
  bool issued = false;
  std::vector<uint64_t> tickets;
};

// One set of weights: a plan and one batch per launch slot.
struct WeightSet {
  uint64_t gen = 0;
  bf_plan *plan = nullptr;
  bool owned = false;           // created by the server (else the caller's plan)
  std::vector<bf_batch *> batches;
  uint64_t live = 0;            // requests of this set submitted and not finished
};

void destroy_set(WeightSet &w) {
  for (bf_batch *b : w.batches)
    if (b) bf_batch_destroy(b);
  w.batches.clear();
  if (w.owned && w.plan) bf_plan_destroy(w.plan);
  w.plan = nullptr;
}

} // namespace

struct bf_server {
  const bf_plan *plan = nullptr;
  bf_server_config cfg{};
  int device = 0;
  uint8_t *arena = nullptr;
  ArenaAllocator alloc;
  AdmissionQueue<uint64_t> *adm = nullptr;
  std::vector<Launch> launches;
  std::deque<WeightSet> sets; // oldest first; back() is current
  std::mutex update_mu;       // serialises bf_server_update_weights
  std::chrono::steady_clock::time_point epoch;

  std::mutex mu;
  std::condition_variable cv_work; // new requests, releases, shutdown
  std::condition_variable cv_done; // finished requests
  std::unordered_map<uint64_t, Request> reqs;
  std::deque<uint64_t> ready;     // admitted, not launched
  std::deque<uint64_t> finished;  // done, not yet returned
  uint64_t next_ticket = 1;
  bool stopping = false;
  std::thread loop;
  std::string loop_error;

  // Pause between taking a launch slot and enqueueing the launch, from the
  // environment variable BF_SERVER_ISSUE_DELAY_US (0 when unset): widens the
  // window in which a launch is taken but not yet enqueued, so that tests can
  // show that the completion scan leaves such launches alone.
  int issue_delay_us = 0;

  uint64_t n_submitted = 0, n_admitted = 0, n_completed = 0, n_released = 0, n_rejected = 0, n_launches = 0;
  uint64_t n_failed = 0, max_batch_seen = 0, batch_requests = 0;

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - epoch).count(); }
};

namespace {

// The weight set of generation gen (nullptr once it is freed).
WeightSet *set_of(bf_server *s, uint64_t gen) {
  for (WeightSet &w : s->sets)
    if (w.gen == gen) return &w;
  return nullptr;
}

// Issues one launch: staging copies of host-fed E, then the kernel, then the
// completion event. Several requests go into one batched launch; a lone
// request takes the plain channel-range launch (no work items to build, and
// the one-beam-tile configuration when it is small). Called without the
// server lock.
int issue(bf_server *s, Launch &L, const bf_plan *plan, bf_batch *batch, const std::vector<bf_chunk> &chunks,
          const std::vector<std::pair<uint8_t *, const Request *>> &copies) {
  for (const auto &c : copies) {
    const hipError_t e =
        hipMemcpyAsync(c.first, c.second->E_src, c.second->e_bytes, hipMemcpyHostToDevice, L.stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync (E staging)");
  }
  int rc;
  if (chunks.size() == 1) {
    const bf_chunk &c = chunks[0];
    rc = bf_form_beams_channels(plan, c.channel, c.nfreq, c.ntimes, c.J, c.E, L.stream);
  } else {
    rc = s->cfg.use_graph ? bf_batch_run_graph(batch, chunks.data(), (int)chunks.size(), L.stream)
                          : bf_batch_run(batch, chunks.data(), (int)chunks.size(), L.stream);
  }
  if (rc) return rc;
  const hipError_t e = hipEventRecord(L.done, L.stream);
  if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
  return BF_OK;
}

// Marks the requests of a launch finished (status rc); lock held.
void finish(bf_server *s, Launch &L, int rc, double t) {
  for (uint64_t id : L.tickets) {
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) continue;
    Request &r = it->second;
    r.t_done = t;
    r.status = rc;
    r.done = true;
    if (r.has_e_piece) { // the staged E is no longer needed
      s->alloc.release(r.e_off);
      r.has_e_piece = false;
    }
    s->adm->finish_slot();
    if (WeightSet *w = set_of(s, r.gen)) --w->live;
    s->finished.push_back(id);
    ++s->n_completed;
    if (rc) ++s->n_failed;
  }
  L.tickets.clear();
  L.busy = false;
  L.issued = false;
}

void admit_locked(bf_server *s) {
  s->adm->admit([&](uint64_t id, const AdmitGrant &g) {
    Request &r = s->reqs[id];
    r.j_off = g.offset[0];
    r.has_j_piece = true;
    if (r.e_on_host) {
      r.e_off = g.offset[1];
      r.has_e_piece = true;
    }
    r.t_admit = s->now();
    s->ready.push_back(id);
    ++s->n_admitted;
  });
}

// Launches admitted requests (one weight set, up to max_batch) on a free
// launch slot; returns false when there is no slot or nothing to launch.
// Called with the lock held; drops it around the launch itself.
bool launch_locked(bf_server *s, std::unique_lock<std::mutex> &lk) {
  int slot = -1;
  for (size_t i = 0; i < s->launches.size(); ++i)
    if (!s->launches[i].busy) {
      slot = (int)i;
      break;
    }
  if (slot < 0 || s->ready.empty() || s->stopping) return false;
  Launch &L = s->launches[slot];
  std::vector<bf_chunk> chunks;
  std::vector<std::pair<uint8_t *, const Request *>> copies;
  const double t = s->now();
  // one weight set per launch: the oldest waiting request's
  const uint64_t gen = s->reqs[s->ready.front()].gen;
  WeightSet *wset = set_of(s, gen); // a set with waiting requests is never freed
  bf_batch *batch = wset->batches[slot];
  const bf_plan *plan = wset->plan;
  while (!s->ready.empty() && (int)chunks.size() < s->cfg.max_batch && s->reqs[s->ready.front()].gen == gen) {
    const uint64_t id = s->ready.front();
    s->ready.pop_front();
    Request &r = s->reqs[id];
    r.t_launch = t;
    const uint8_t *E = r.E_src;
    if (r.e_on_host) {
      E = s->arena + r.e_off;
      copies.emplace_back(s->arena + r.e_off, &r);
    }
    chunks.push_back(bf_chunk{E, s->arena + r.j_off, r.ntimes, r.channel, r.nfreq});
    L.tickets.push_back(id);
  }
  for (uint64_t id : L.tickets) s->reqs[id].batch_size = (int)L.tickets.size();
  L.busy = true; // the slot is ours while the lock is dropped
  L.issued = false;
  ++s->n_launches;
  s->batch_requests += L.tickets.size();
  s->max_batch_seen = std::max<uint64_t>(s->max_batch_seen, L.tickets.size());
  lk.unlock();
  if (s->issue_delay_us > 0) std::this_thread::sleep_for(std::chrono::microseconds(s->issue_delay_us));
  const int rc = issue(s, L, plan, batch, chunks, copies); // Request pointers stay valid: entries are only erased on release
  if (rc) (void)hipStreamSynchronize(L.stream); // no staging copy of a failed launch may still write the arena
  lk.lock();
  L.issued = true; // from here on `done` belongs to this launch
  if (rc) {
    s->loop_error = g_last_error;
    finish(s, L, rc, s->now());
    s->cv_done.notify_all();
  }
  return true;
}

void run_loop(bf_server *s) {
  if (hipSetDevice(s->device) != hipSuccess) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->loop_error = "hipSetDevice failed in the server thread";
  }
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    bool progressed = false;
    if (!s->stopping) admit_locked(s); // shutting down: only finish what is on the device

    if (launch_locked(s, lk)) progressed = true;

    // Completion.
    bool any_busy = false;
    for (Launch &L : s->launches) {
      if (!L.busy) continue;
      if (!L.issued) { // another thread is still enqueueing this launch
        any_busy = true;
        continue;
      }
      const hipError_t q = hipEventQuery(L.done);
      if (q == hipErrorNotReady) {
        any_busy = true;
        continue;
      }
      const int rc = q == hipSuccess ? BF_OK : hip_fail(q, "launch");
      if (rc) s->loop_error = g_last_error;
      finish(s, L, rc, s->now());
      s->cv_done.notify_all();
      progressed = true;
    }

    if (s->stopping && !any_busy) break;
    if (progressed) continue;
    if (any_busy) {
      if (s->cfg.poll_us > 0) {
        s->cv_work.wait_for(lk, std::chrono::microseconds(s->cfg.poll_us));
      } else {
        // Poll again at once: a timed wait can add the OS's wake-up delay
        // (measured 45 us on one host) to every request's latency. The lock
        // is dropped so that submitters and pollers get in.
        lk.unlock();
        std::this_thread::yield();
        lk.lock();
      }
    } else {
      s->cv_work.wait(lk);
    }
  }
}

void fill_result(const bf_server *s, const Request &r, bf_result *out) {
  std::memset(out, 0, sizeof *out);
  out->ticket = r.ticket;
  out->tag = r.tag;
  out->J = s->arena + r.j_off;
  out->j_bytes = r.j_bytes;
  out->status = r.status;
  out->batch_size = r.batch_size;
  out->t_submit = r.t_submit;
  out->t_admit = r.t_admit;
  out->t_launch = r.t_launch;
  out->t_done = r.t_done;
}

int pop_result(bf_server *s, bf_result *out) {
  const uint64_t id = s->finished.front();
  s->finished.pop_front();
  fill_result(s, s->reqs[id], out);
  return BF_OK;
}

} // namespace

extern "C" {

int bf_server_create(const bf_plan *p, const bf_server_config *cfg, bf_server **out) {
  if (!out) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (!p || !cfg) return fail(BF_ERR_INVALID_ARGUMENT, "NULL plan or config");
  if (cfg->hbm_budget == 0) return fail(BF_ERR_INVALID_ARGUMENT, "hbm_budget must be > 0");
  if (cfg->max_inflight < 0 || cfg->max_batch < 0 || cfg->max_launches < 0 || cfg->poll_us < 0 ||
      cfg->poll_us > 1000000)
    return fail(BF_ERR_INVALID_ARGUMENT, "negative limit, or poll_us outside 0..1000000");
  bf_server *s = new bf_server();
  s->plan = p;
  s->cfg = *cfg;
  if (!s->cfg.max_inflight) s->cfg.max_inflight = 256;
  if (!s->cfg.max_batch) s->cfg.max_batch = 128;
  if (!s->cfg.max_launches) s->cfg.max_launches = 2;
  s->device = p->device;
  s->epoch = std::chrono::steady_clock::now();
  if (const char *d = std::getenv("BF_SERVER_ISSUE_DELAY_US")) s->issue_delay_us = std::max(0, std::min(1000000, std::atoi(d)));
  auto bail = [&](int rc) {
    bf_server_destroy(s);
    return rc;
  };
  hipError_t e;
  if ((e = hipSetDevice(s->device)) != hipSuccess) return bail(hip_fail(e, "hipSetDevice"));
  if ((e = hipMalloc(&s->arena, cfg->hbm_budget)) != hipSuccess) return bail(hip_fail(e, "hipMalloc (HBM budget)"));
  s->alloc.reset(cfg->hbm_budget, 256);
  s->adm = new AdmissionQueue<uint64_t>(&s->alloc, s->cfg.max_inflight);
  s->launches.resize(s->cfg.max_launches);
  s->sets.emplace_back();
  s->sets.back().plan = const_cast<bf_plan *>(p); // not owned: never modified or freed here
  for (Launch &L : s->launches) {
    bf_batch *b = nullptr;
    if (int rc = bf_batch_create(p, s->cfg.max_batch, &b)) return bail(rc);
    s->sets.back().batches.push_back(b);
    if ((e = hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking)) != hipSuccess)
      return bail(hip_fail(e, "hipStreamCreate"));
    if ((e = hipEventCreateWithFlags(&L.done, hipEventDisableTiming)) != hipSuccess)
      return bail(hip_fail(e, "hipEventCreate"));
  }
  s->loop = std::thread(run_loop, s);
  *out = s;
  return BF_OK;
}

int bf_server_destroy(bf_server *s) {
  if (!s) return BF_OK;
  if (s->loop.joinable()) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stopping = true;
    }
    s->cv_work.notify_all();
    s->loop.join(); // returns once no launch is in flight
  }
  for (Launch &L : s->launches)
    if (L.stream) (void)hipStreamSynchronize(L.stream);
  for (WeightSet &w : s->sets) destroy_set(w);
  s->sets.clear();
  for (Launch &L : s->launches) {
    if (L.done) (void)hipEventDestroy(L.done);
    if (L.stream) (void)hipStreamDestroy(L.stream);
  }
  delete s->adm;
  if (s->arena) (void)hipFree(s->arena);
  delete s;
  return BF_OK;
}

int bf_server_submit_many(bf_server *s, const bf_request *reqs, int n, uint64_t *tickets) {
  if (!s || n < 0 || (n > 0 && !reqs)) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument or n < 0");
  const bf_shape &sh = s->plan->shape;
  std::vector<Request> rs((size_t)n);
  std::vector<AdmitNeed> needs((size_t)n);
  std::vector<size_t> totals((size_t)n);
  for (int i = 0; i < n; ++i) { // validate everything before queueing anything
    const bf_request *req = &reqs[i];
    const int nf = req->nfreq == 0 ? 1 : req->nfreq;
    // ntimes <= kMaxTimes (the batch kernel's limit) keeps the byte counts below in int64
    if (!req->E || req->ntimes < 1 || req->ntimes > kMaxTimes || nf < 1 || nf > 65535 || req->channel < 0 ||
        (int64_t)req->channel + nf > sh.nfrequencies || req->reserved != 0 ||
        (req->e_on_host != 0 && req->e_on_host != 1))
      return fail(BF_ERR_INVALID_ARGUMENT, "bad request " + std::to_string(i) +
                                               " (NULL E, ntimes < 1, channels outside the table, flags)");
    Request &r = rs[i];
    r.tag = req->tag;
    r.E_src = req->E;
    r.e_on_host = req->e_on_host != 0;
    r.ntimes = req->ntimes;
    r.channel = req->channel;
    r.nfreq = nf;
    r.j_bytes = (size_t)(sh.nbeams * nf * sh.npolarizations * req->ntimes);
    r.e_bytes = (size_t)(req->ntimes * nf * sh.ndishes * sh.npolarizations);
    needs[i].bytes[0] = r.j_bytes;
    needs[i].bytes[1] = r.e_on_host ? r.e_bytes : 0;
    totals[i] = s->alloc.round(needs[i].bytes[0]) + (needs[i].bytes[1] ? s->alloc.round(needs[i].bytes[1]) : 0);
  }
  {
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->stopping) return fail(BF_ERR_CLOSED, "server is shutting down");
    for (int i = 0; i < n; ++i)
      if (totals[i] > s->alloc.capacity()) {
        ++s->n_rejected;
        return fail(BF_ERR_OVER_BUDGET, "request " + std::to_string(i) + " needs " + std::to_string(totals[i]) +
                                            " bytes, budget is " + std::to_string(s->alloc.capacity()));
      }
    const double t = s->now();
    WeightSet &cur = s->sets.back();
    for (int i = 0; i < n; ++i) {
      Request &r = rs[i];
      r.ticket = s->next_ticket++;
      r.t_submit = t;
      r.gen = cur.gen;
      ++cur.live;
      s->reqs[r.ticket] = r;
      s->adm->push(r.ticket, needs[i]);
      ++s->n_submitted;
      if (tickets) tickets[i] = r.ticket;
    }
    if (s->cfg.poll_us == 0) {
      // Latency first: admit and launch from the submitting thread when a
      // launch slot is free, instead of waiting for the loop thread to wake
      // (about 5-10 us); the loop then polls the launch's completion.
      admit_locked(s);
      int prev = -1;
      if (hipGetDevice(&prev) == hipSuccess && (prev == s->device || hipSetDevice(s->device) == hipSuccess)) {
        launch_locked(s, lk);
        if (prev != s->device) (void)hipSetDevice(prev); // leave the caller's current device as it was
      }
    }
  }
  s->cv_work.notify_all();
  return BF_OK;
}

int bf_server_submit(bf_server *s, const bf_request *req, uint64_t *ticket) {
  if (!req) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  return bf_server_submit_many(s, req, 1, ticket);
}

int bf_server_poll(bf_server *s, bf_result *out) {
  if (!s || !out) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->finished.empty()) return BF_ERR_NOT_READY;
  return pop_result(s, out);
}

int bf_server_wait(bf_server *s, double timeout_s, bf_result *out) {
  if (!s || !out) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  const auto pred = [&] { return !s->finished.empty(); };
  if (timeout_s < 0)
    s->cv_done.wait(lk, pred);
  else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
    return BF_ERR_NOT_READY;
  return pop_result(s, out);
}

int bf_server_read(bf_server *s, uint64_t ticket, uint8_t *host, size_t len) {
  if (!s || !host) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  uint8_t *src = nullptr;
  size_t n = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end() || !it->second.done) return fail(BF_ERR_INVALID_ARGUMENT, "unknown or unfinished ticket");
    src = s->arena + it->second.j_off;
    n = std::min(len, it->second.j_bytes);
  }
  const hipError_t e = hipMemcpy(host, src, n, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy (result)");
  return BF_OK;
}

int bf_server_release(bf_server *s, uint64_t ticket) {
  if (!s) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end() || !it->second.done) return fail(BF_ERR_INVALID_ARGUMENT, "unknown or unfinished ticket");
    Request &r = it->second;
    if (r.has_j_piece) s->alloc.release(r.j_off);
    if (r.has_e_piece) s->alloc.release(r.e_off);
    // drop it from the not-yet-returned list too, if the caller never polled it
    auto f = std::find(s->finished.begin(), s->finished.end(), ticket);
    if (f != s->finished.end()) s->finished.erase(f);
    s->reqs.erase(it);
    ++s->n_released;
  }
  s->cv_work.notify_all();
  return BF_OK;
}

int bf_server_stats(bf_server *s, bf_server_counters *out) {
  if (!s || !out) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof *out);
  out->submitted = s->n_submitted;
  out->admitted = s->n_admitted;
  out->completed = s->n_completed;
  out->released = s->n_released;
  out->rejected = s->n_rejected;
  out->failed = s->n_failed;
  out->launches = s->n_launches;
  out->max_batch = s->max_batch_seen;
  out->waiting = s->adm->waiting();
  out->inflight = (uint64_t)s->adm->slots_used();
  out->arena_bytes = s->alloc.capacity();
  out->bytes_in_use = s->alloc.in_use();
  out->peak_bytes = s->alloc.peak();
  out->mean_batch = s->n_launches ? (double)s->batch_requests / (double)s->n_launches : 0.0;
  out->clock = s->now();
  return BF_OK;
}

double bf_server_clock(const bf_server *s) { return s ? s->now() : 0.0; }

int bf_server_update_weights(bf_server *s, const uint8_t *A, const float *G) {
  if (!s || !A || !G) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> ulk(s->update_mu);
  hipError_t e = hipSetDevice(s->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  WeightSet w;
  w.owned = true;
  // a plan of the same shape: expands the new weights on the device while the
  // loop keeps launching with the current set
  if (int rc = bf_plan_create(&s->plan->shape, A, G, &w.plan)) return rc;
  for (size_t i = 0; i < s->launches.size(); ++i) {
    bf_batch *b = nullptr;
    if (int rc = bf_batch_create(w.plan, s->cfg.max_batch, &b)) {
      destroy_set(w);
      return rc;
    }
    w.batches.push_back(b);
  }
  std::vector<WeightSet> retired;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) {
      destroy_set(w);
      return fail(BF_ERR_CLOSED, "server is shutting down");
    }
    w.gen = s->sets.back().gen + 1;
    s->sets.push_back(std::move(w));
    // older sets without pending requests are no longer referenced
    for (auto it = s->sets.begin(); it + 1 != s->sets.end();) {
      if (it->live == 0) {
        retired.push_back(std::move(*it));
        it = s->sets.erase(it);
      } else {
        ++it;
      }
    }
  }
  for (WeightSet &r : retired) destroy_set(r);
  return BF_OK;
}

} // extern "C"

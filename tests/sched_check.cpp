// Host checks of the serving loop's device-independent policy
// (beamforming_amd/csrc/bf_sched.hpp): arena sub-allocation and FIFO
// admission. Built with g++ and run by tests/test_sched.py; prints "ok".
#include "bf_sched.hpp"

#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <vector>

#define CHECK(c)                                                                                                       \
  do {                                                                                                                 \
    if (!(c)) {                                                                                                        \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);                                                       \
      std::exit(1);                                                                                                    \
    }                                                                                                                  \
  } while (0)

using bf::AdmissionQueue;
using bf::AdmitGrant;
using bf::AdmitNeed;
using bf::ArenaAllocator;

static void arena_basics() {
  ArenaAllocator a(1000, 256); // capacity rounds down to 768
  CHECK(a.capacity() == 768);
  size_t o1, o2, o3, o4;
  CHECK(a.alloc(1, &o1) && o1 == 0);
  CHECK(a.alloc(256, &o2) && o2 == 256);
  CHECK(a.alloc(200, &o3) && o3 == 512);
  CHECK(!a.alloc(1, &o4)); // full
  CHECK(a.in_use() == 768 && a.peak() == 768);
  CHECK(a.release(o2));
  CHECK(!a.release(o2)); // double free is refused
  CHECK(!a.release(12345));
  CHECK(!a.alloc(257, &o4)); // only 256 free
  CHECK(a.alloc(256, &o4) && o4 == 256);
  CHECK(a.release(o1) && a.release(o3) && a.release(o4));
  CHECK(a.in_use() == 0 && a.free_ranges() == 1 && a.largest_free() == 768);
}

// Random alloc/free against a byte map: no overlap, exact accounting, full
// coalescing at the end.
static void arena_random() {
  std::mt19937_64 rng(7);
  const size_t cap = 256 * 1000;
  ArenaAllocator a(cap, 256);
  std::vector<int> owner(cap / 256, -1);
  std::vector<std::pair<size_t, size_t>> live; // offset, rounded size
  size_t used = 0;
  for (int step = 0; step < 20000; ++step) {
    const bool do_alloc = live.empty() || (rng() % 3 != 0);
    if (do_alloc) {
      const size_t bytes = 1 + rng() % (256 * 40);
      size_t off;
      if (a.alloc(bytes, &off)) {
        const size_t len = a.round(bytes);
        CHECK(off % 256 == 0 && off + len <= cap);
        for (size_t b = off / 256; b < (off + len) / 256; ++b) {
          CHECK(owner[b] == -1);
          owner[b] = step;
        }
        live.push_back({off, len});
        used += len;
      } else {
        // refusal is only allowed when no free run of that size exists
        const size_t need = a.round(bytes) / 256;
        size_t run = 0, best = 0;
        for (int o : owner) {
          run = o == -1 ? run + 1 : 0;
          best = run > best ? run : best;
        }
        CHECK(best < need);
      }
    } else {
      const size_t k = rng() % live.size();
      CHECK(a.release(live[k].first));
      for (size_t b = live[k].first / 256; b < (live[k].first + live[k].second) / 256; ++b) owner[b] = -1;
      used -= live[k].second;
      live[k] = live.back();
      live.pop_back();
    }
    CHECK(a.in_use() == used);
  }
  for (auto &l : live) CHECK(a.release(l.first));
  CHECK(a.in_use() == 0 && a.free_ranges() == 1 && a.largest_free() == cap);
}

// FIFO admission: the head blocks younger requests until it fits; slots cap
// the number admitted; a failed two-piece grant leaves the arena unchanged.
static void admission() {
  ArenaAllocator a(256 * 10, 256);
  AdmissionQueue<int> q(&a, 3);
  auto need = [](size_t x, size_t y = 0) {
    AdmitNeed n;
    n.bytes[0] = x;
    n.bytes[1] = y;
    return n;
  };
  std::vector<int> order;
  std::vector<AdmitGrant> grants(16);
  auto grant = [&](int id, const AdmitGrant &g) {
    order.push_back(id);
    grants[id] = g;
  };
  q.push(0, need(256 * 4));
  q.push(1, need(256 * 4));
  q.push(2, need(256 * 4)); // does not fit after 0 and 1 (2 blocks left)
  q.push(3, need(256));     // would fit, but must not overtake 2
  CHECK(q.admit(grant) == 2);
  CHECK(order == std::vector<int>({0, 1}));
  CHECK(q.waiting() == 2 && q.slots_used() == 2);
  // free request 0's memory (finished and released)
  q.finish_slot();
  CHECK(a.release(grants[0].offset[0]));
  CHECK(q.admit(grant) == 2); // 2 fits now, then 3
  CHECK(order == std::vector<int>({0, 1, 2, 3}));
  CHECK(q.slots_used() == 3);
  // slot cap: memory is free but all 3 slots are taken
  q.push(5, need(1));
  CHECK(q.admit(grant) == 0);
  q.finish_slot();
  CHECK(q.admit(grant) == 1);
  CHECK(order.back() == 5);
  // two pieces, second does not fit: nothing stays allocated
  const size_t before = a.in_use();
  q.push(4, need(256, 256 * 9));
  q.finish_slot();
  CHECK(q.admit(grant) == 0);
  CHECK(a.in_use() == before);
  CHECK(q.waiting() == 1);
}

static void percentiles() {
  std::vector<double> v;
  for (int i = 100; i >= 1; --i) v.push_back(i);
  CHECK(bf::percentile(v, 0.5) == 50);
  CHECK(bf::percentile(v, 0.99) == 99);
  CHECK(bf::percentile(v, 1.0) == 100);
  CHECK(bf::percentile(v, 0.0) == 1);
  std::vector<double> e;
  CHECK(bf::percentile(e, 0.5) == 0);
}

int main() {
  arena_basics();
  arena_random();
  admission();
  percentiles();
  std::printf("ok\n");
  return 0;
}

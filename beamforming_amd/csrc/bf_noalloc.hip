// bf_noalloc.hip -- bf_adler32_device_noalloc: the variant of
// bf_adler32_device that allocates nothing, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses
// that forbid allocation).
//
// Its scratch is this file's static device array `scratch`: the 2 partial
// words of each workgroup of the partial pass, then the 32-bit result. Every
// GPU the library runs on holds its own copy until the process ends, at its
// own address, found once per device with hipGetSymbolAddress on the first
// call on that device (with it current). Host code only ever passes that
// looked-up address, never the array's name. Nothing in the scratch needs to
// start at zero: each call writes every partial pair it reads and the result.
// Every call on one GPU shares that GPU's scratch, so calls that overlap on
// one GPU (on different streams or threads) race; calls on one stream one
// after the other are safe.
#include "beamform.h"
#include "bf_plan.hpp"

#include <mutex>

using namespace bf;

// kAdlerScratchWords = 2 * kAdlerBlocks + 2 words = 32 KiB + 16 bytes.
static __device__ __align__(256) unsigned long long scratch[kAdlerScratchWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned long long *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device.
int scratch_ptr(unsigned long long **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(BF_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(BF_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = static_cast<unsigned long long *>(p);
  }
  *out = g_scratch_ptr[dev];
  return BF_OK;
}

} // namespace

extern "C" int bf_adler32_device_noalloc(const uint8_t *data, size_t len, uint32_t *out, void *stream) {
  if (!data || !out) return fail(BF_ERR_INVALID_ARGUMENT, "NULL argument");
  // The scratch holds the partials of kAdlerBlocks workgroups. The partial
  // pass never runs more (each folds any number of 16-byte vectors), so no
  // length is too long; the check guards that invariant.
  if (adler32_blocks(data, (uint64_t)len) > kAdlerBlocks)
    return fail(BF_ERR_UNSUPPORTED, "adler32: more partial sums than the static scratch holds");
  unsigned long long *ptr = nullptr;
  if (int rc = scratch_ptr(&ptr)) return rc;
  return adler32_run(data, len, ptr, out, reinterpret_cast<hipStream_t>(stream));
}

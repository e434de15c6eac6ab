// ref_harness.cxx -- drives the reference's own CPU path (cpu.cxx, arraysizes.cxx,
// adler32.c, compiled from /root/reference by oracle/ref.mk) to produce golden
// vectors for the oracle. TEST INFRASTRUCTURE ONLY.
//
// The reference fixes its shape at compile time (arraysizes.hxx:22-27). This
// harness runs at that shape. It calls the reference's form_beams() and
// adler32(); inputs come either from the reference's setup() or from the
// oracle's seeded generator (bf_oracle.c), so the oracle can be compared with
// the reference on random data.
//
// Usage:
//   ref_harness shape
//   ref_harness codec
//   ref_harness adler32 <seed> <len>
//   ref_harness pattern [nsamples]
//   ref_harness seeded <seed> <nsamples>
//   ref_harness time      (wall time of one form_beams call on setup() inputs)
// Each prints one JSON object on stdout.
#include "adler32.h"
#include "arraysizes.hxx"
#include "icomplex4.hxx"

#include "bf_oracle.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

void form_beams(ucomplex4 *restrict const Jarray, const ucomplex4 *restrict const Earray,
                const ucomplex4 *restrict const Aarray, const float *restrict Garray);

static void print_shape() {
  std::printf("\"shape\": {\"ntimes\": %zu, \"nfrequencies\": %zu, \"ndishes\": %zu, \"npolarizations\": %zu, "
              "\"nbeams\": %zu}",
              ntimes, nfrequencies, ndishes, npolarizations, nbeams);
}

static void print_samples(const std::vector<ucomplex4> &J, size_t nsamples) {
  // Sample positions: SplitMix64 stream with seed 0x5EED, reduced modulo |J|.
  std::printf("\"samples\": [");
  for (size_t i = 0; i < nsamples; ++i) {
    const uint64_t pos = bfo_splitmix64_at(0x5EEDull, i) % J.size();
    std::printf("%s[%llu, %u]", i ? ", " : "", (unsigned long long)pos, (unsigned)J[pos].data);
  }
  std::printf("]");
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: ref_harness shape|codec|adler32|pattern|seeded ...\n");
    return 2;
  }
  const char *mode = argv[1];
  if (!std::strcmp(mode, "shape")) {
    std::printf("{");
    print_shape();
    std::printf("}\n");
    return 0;
  }
  if (!std::strcmp(mode, "codec")) {
    // Decode table for all 256 bytes and encode table for re, im in [-8, 7].
    std::printf("{\"decode\": [");
    for (int x = 0; x < 256; ++x) {
      ucomplex4 c;
      c.data = (unsigned char)x;
      std::printf("%s[%d, %d]", x ? ", " : "", (int)c.real(), (int)c.imag());
    }
    std::printf("], \"encode\": [");
    bool first = true;
    for (int re = -8; re <= 7; ++re)
      for (int im = -8; im <= 7; ++im) {
        std::printf("%s[%d, %d, %d]", first ? "" : ", ", re, im, (int)ucomplex4(re, im).data);
        first = false;
      }
    std::printf("]}\n");
    return 0;
  }
  if (!std::strcmp(mode, "adler32")) {
    if (argc < 4) return 2;
    const uint64_t seed = std::strtoull(argv[2], nullptr, 0);
    const size_t len = std::strtoull(argv[3], nullptr, 0);
    std::vector<unsigned char> buf(len);
    bfo_fill_uniform(buf.data(), len, seed, 0);
    std::printf("{\"seed\": %llu, \"len\": %zu, \"adler32\": %u}\n", (unsigned long long)seed, len,
                (unsigned)adler32(buf.data(), len));
    return 0;
  }
  if (!std::strcmp(mode, "pattern") || !std::strcmp(mode, "seeded")) {
    std::vector<ucomplex4> E, A, J;
    std::vector<float> G;
    setup(E, A, G, J); // reference inputs (arraysizes.cxx:8-36)
    uint64_t seed = 0;
    size_t nsamples = 0;
    if (!std::strcmp(mode, "pattern") && argc >= 3) nsamples = std::strtoull(argv[2], nullptr, 0);
    if (!std::strcmp(mode, "seeded")) {
      if (argc < 4) return 2;
      seed = std::strtoull(argv[2], nullptr, 0);
      nsamples = std::strtoull(argv[3], nullptr, 0);
      // Oracle generator for E and A; G stays the reference formula.
      bfo_fill_uniform(reinterpret_cast<uint8_t *>(E.data()), E.size(), seed, 0);
      bfo_fill_uniform(reinterpret_cast<uint8_t *>(A.data()), A.size(), seed ^ 0xA5A5A5A5A5A5A5A5ull, 0);
    }
    std::memset(J.data(), 0, J.size());
    form_beams(J.data(), E.data(), A.data(), G.data()); // reference cpu.cxx:15-46
    const uint32_t ck = adler32(reinterpret_cast<const unsigned char *>(J.data()), J.size());
    std::printf("{");
    print_shape();
    std::printf(", \"mode\": \"%s\", \"seed\": %llu, \"adler32\": %u, ", mode, (unsigned long long)seed, (unsigned)ck);
    print_samples(J, nsamples);
    std::printf("}\n");
    return 0;
  }
  if (!std::strcmp(mode, "time")) {
    // the reference's timing flow: wall clock around form_beams only (cpu.cxx:58-61)
    std::vector<ucomplex4> E, A, J;
    std::vector<float> G;
    setup(E, A, G, J);
    const auto t0 = std::chrono::steady_clock::now();
    form_beams(J.data(), E.data(), A.data(), G.data());
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const uint32_t ck = adler32(reinterpret_cast<const unsigned char *>(J.data()), J.size());
    std::printf("{");
    print_shape();
    std::printf(", \"mode\": \"time\", \"seconds\": %.6f, \"adler32\": %u}\n", dt, (unsigned)ck);
    return 0;
  }
  std::fprintf(stderr, "unknown mode %s\n", mode);
  return 2;
}

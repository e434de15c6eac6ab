// bf_form_beams -- the reference driver flow (cpu.cxx:48-67, cuda.cu:264-322)
// on the MI355X library: setup() inputs -> form_beams -> adler32 check.
//
//   bf_form_beams [ntimes nfrequencies ndishes nbeams] [--layout bfpt|bftp]
//                 [--expect 0xXXXXXXXX] [--repeat N]
//
// Without sizes it runs the reference's shipped configuration
// (arraysizes.hxx:22-27). The checksum is compared with the reference's table
// (arraysizes.cxx:54-59) or with --expect; a mismatch exits with status 1 as
// in arraysizes.cxx:82-86.
#include "beamform.hpp"

#include <sys/time.h>

#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using bf::ucomplex4;

static double gettime() {
  timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec / 1.0e+6;
}

// arraysizes.cxx:16-32: E[n] = A[n] = (n % 15 - 7, (n + 1) % 15 - 7) over the
// byte index; G[n] = float(n) / Gsize / ndishes * (15 + n % 15) / 30.
static void setup(std::vector<ucomplex4> &E, std::vector<ucomplex4> &A, std::vector<float> &G,
                  std::vector<ucomplex4> &J) {
  std::printf("Setting up input data...\n");
  E.resize(bf::e_count());
  A.resize(bf::a_count());
  G.resize(bf::g_count());
  J.resize(bf::j_count());
  for (size_t n = 0; n < E.size(); ++n) E[n] = ucomplex4((signed char)(n % 15) - 7, (signed char)((n + 1) % 15) - 7);
  for (size_t n = 0; n < A.size(); ++n) A[n] = ucomplex4((signed char)(n % 15) - 7, (signed char)((n + 1) % 15) - 7);
  const float ng = (float)G.size(), nd = (float)bf::current_shape().ndishes;
  for (size_t n = 0; n < G.size(); ++n) G[n] = (float)n / ng / nd * (float)(15 + n % 15) / 30.0f;
}

struct Known {
  int64_t T, F, D, B;
  uint32_t checksum;
};
// adler32 of J ([b][f][p][t]) for setup() inputs. The first six are the
// reference's table (arraysizes.cxx:54-59; the F = 1 rows correspond to the
// configurations at arraysizes.hxx:58-70); the others were computed with the
// reference cpu.cxx at those sizes.
static const Known kKnown[] = {
    {32768, 16, 512, 96, 0x27303ea9}, {32768, 32, 512, 128, 0x4de6498f}, {32768, 32, 512, 96, 0xc4fb6a4c},
    {32, 32, 512, 128, 0xab69dfee},   {8, 1, 32, 8, 0x3a3344bc},          {1, 1, 1, 1, 0x019a0111},
    {256, 4, 64, 16, 0x07cee085},     {8192, 16, 512, 512, 0x63134785},   {8192, 68, 512, 96, 0xee913082},
    {32, 32, 512, 96, 0x691be8a7},
};

int main(int argc, char **argv) {
  std::printf("beamforming.mi355x\n");
  int64_t dims[4] = {8192, 68, 512, 96};
  int ndims = 0;
  int32_t layout = BF_J_BFPT;
  bool have_expect = false;
  uint32_t expect = 0;
  int repeat = 1;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--layout") && i + 1 < argc) {
      layout = !std::strcmp(argv[++i], "bftp") ? BF_J_BFTP : BF_J_BFPT;
    } else if (!std::strcmp(argv[i], "--expect") && i + 1 < argc) {
      expect = (uint32_t)std::strtoul(argv[++i], nullptr, 0);
      have_expect = true;
    } else if (!std::strcmp(argv[i], "--repeat") && i + 1 < argc) {
      repeat = std::atoi(argv[++i]);
    } else if (ndims < 4) {
      dims[ndims++] = std::strtoll(argv[i], nullptr, 0);
    } else {
      std::fprintf(stderr, "usage: %s [ntimes nfrequencies ndishes nbeams] [--layout bfpt|bftp] "
                           "[--expect 0xCHECKSUM] [--repeat N]\n", argv[0]);
      return 2;
    }
  }
  if (ndims != 0 && ndims != 4) {
    std::fprintf(stderr, "give all four sizes or none\n");
    return 2;
  }
  bf::set_shape(dims[0], dims[1], dims[2], dims[3], 2, layout);
  std::printf("ntimes=%" PRId64 " nfrequencies=%" PRId64 " ndishes=%" PRId64 " npolarizations=2 nbeams=%" PRId64 "\n",
              dims[0], dims[1], dims[2], dims[3]);

  std::vector<ucomplex4> E, A, J;
  std::vector<float> G;
  setup(E, A, G, J);

  std::printf("Forming beams...\n");
  double best = 1e30;
  try {
    for (int r = 0; r < repeat; ++r) {
      const double t0 = gettime();
      bf::form_beams(J.data(), E.data(), A.data(), G.data());
      const double t1 = gettime();
      best = t1 - t0 < best ? t1 - t0 : best;
    }
  } catch (const bf::error &e) {
    std::printf("ERROR -- %s\n", e.what());
    return 1;
  }
  std::printf("Elapsed time: %g seconds (host arrays: plan, copies and kernel)\n", best);

  std::printf("Calculating checksum...\n");
  const uint32_t checksum = bf_adler32(reinterpret_cast<const uint8_t *>(J.data()), J.size());
  std::printf("Checksum: 0x%08x\n", checksum);
  if (!have_expect && layout == BF_J_BFPT) {
    for (const Known &k : kKnown)
      if (k.T == dims[0] && k.F == dims[1] && k.D == dims[2] && k.B == dims[3]) {
        expect = k.checksum;
        have_expect = true;
      }
  }
  if (!have_expect) {
    std::printf("No reference checksum for this shape.\nDone.\n");
    return 0;
  }
  if (checksum != expect) {
    std::printf("Expected: 0x%08x\n", expect);
    std::printf("ERROR -- CHECKSUM MISMATCH\n");
    return 1;
  }
  std::printf("Checksum matches.\n");
  std::printf("Done.\n");
  return 0;
}

// Host checks of the per-lane arithmetic of the gfx950 kernels
// (beamforming_amd/csrc/bf_device.hpp), compiled for the CPU: the byte
// permute, nibble unpack, operand windows, weight packing (int8 and fp8),
// the fp8 reading of the operands, nibble-offset
// corrections, the fast and exact requantisation and the output packing,
// each against the oracle's codec (icomplex4.hxx:22-40), its requantisation
// (cpu.cxx:37-39) and the textbook complex product (cpu.cxx:24-35).
// Prints "ok" and exits 0 when every check passes.
#include "bf_device.hpp"

extern "C" {
#include "bf_oracle.h"
}

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>

using namespace bf;

namespace {

int failures = 0;

void expect(bool ok, const char *what, long long a, long long b) {
  if (!ok && failures++ < 20) std::printf("FAIL %s: %lld vs %lld\n", what, a, b);
}

uint64_t state = 0x9e3779b97f4a7c15ull;
uint64_t next() {
  uint64_t z = (state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

int8_t byte_of(const v4i v, int i) {
  const int w[4] = {v.x, v.y, v.z, v.w};
  return (int8_t)(uint8_t)((uint32_t)w[i >> 2] >> (8 * (i & 3)));
}

int dot16(const v4i a, const v4i b) {
  int s = 0;
  for (int i = 0; i < 16; ++i) s += (int)byte_of(a, i) * (int)byte_of(b, i);
  return s;
}

void check_codec() {
  for (int x = 0; x < 256; ++x) {
    expect(dec_re(x) == bfo_dec_re((uint8_t)x), "dec_re", dec_re(x), bfo_dec_re((uint8_t)x));
    expect(dec_im(x) == bfo_dec_im((uint8_t)x), "dec_im", dec_im(x), bfo_dec_im((uint8_t)x));
  }
  for (int re = -8; re <= 7; ++re)
    for (int im = -8; im <= 7; ++im) expect(enc(re, im) == bfo_enc(re, im), "enc", enc(re, im), bfo_enc(re, im));
}

void check_perm() {
  // selector bytes 0-7 pick bytes of hi:lo, 12 gives zero
  const uint32_t hi = 0x77665544u, lo = 0x33221100u;
  expect(perm_bytes(hi, lo, 0x07050301u) == 0x77553311u, "perm odd", perm_bytes(hi, lo, 0x07050301u), 0x77553311u);
  expect(perm_bytes(hi, lo, 0x06040200u) == 0x66442200u, "perm even", perm_bytes(hi, lo, 0x06040200u), 0x66442200u);
  expect(perm_bytes(hi, lo, 0x0c0c0400u) == 0x00004400u, "perm zero", perm_bytes(hi, lo, 0x0c0c0400u), 0x00004400u);
}

// One K step of one lane: 8 dishes x 2 polarizations of E against the
// weights of one beam, both operand variants, for each polarization.
void check_operands() {
  for (int trial = 0; trial < 20000; ++trial) {
    uint8_t e[16], a[8];
    for (int i = 0; i < 16; ++i) e[i] = (uint8_t)next();
    for (int i = 0; i < 8; ++i) a[i] = (uint8_t)next();
    if (trial < 256) { // every E byte value, extreme weights
      std::memset(e, trial, sizeof e);
      std::memset(a, trial & 1 ? 0x00 : 0xff, sizeof a); // -8-8i and 7+7i
    }
    uint32_t w[4];
    std::memcpy(w, e, 16);
    const v4i ev{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
    const v4i W = pack_weights(a);
    int sum_ar = 0, sum_ai = 0;
    for (int d = 0; d < 8; ++d) {
      sum_ar += bfo_dec_re(a[d]);
      sum_ai += bfo_dec_im(a[d]);
    }
    int corr_re, corr_im;
    corrections(sum_ar, sum_ai, &corr_re, &corr_im);
    for (int p = 0; p < 2; ++p) {
      const v8i u = unpack_e_pol(ev, p ? 0x07050301u : 0x06040200u);
      int want_re = 0, want_im = 0;
      for (int d = 0; d < 8; ++d) {
        const int er = bfo_dec_re(e[2 * d + p]), ei = bfo_dec_im(e[2 * d + p]);
        const int ar = bfo_dec_re(a[d]), ai = bfo_dec_im(a[d]);
        want_re += ar * er - ai * ei;
        want_im += ar * ei + ai * er;
      }
      const int got_re = dot16(op_v1(u), W) - corr_re;
      const int got_im = dot16(op_v2(u), W) - corr_im;
      expect(got_re == want_re, "raw re", got_re, want_re);
      expect(got_im == want_im, "raw im", got_im, want_im);
    }
  }
}

// OCP 6-bit float formats (format codes 2 and 3 of v_mfma_scale_f32_16x16x128_f8f6f4): E2M3
// (exponent bias 1, 3 mantissa bits) and E3M2 (bias 3, 2 mantissa bits);
// neither has infinities or NaNs.
double dec_e2m3(uint32_t x) {
  const int s = (x >> 5) & 1, e = (x >> 3) & 3, m = x & 7;
  const double v = e == 0 ? m / 8.0 : std::ldexp(1.0 + m / 8.0, e - 1);
  return s ? -v : v;
}
double dec_e3m2(uint32_t x) {
  const int s = (x >> 5) & 1, e = (x >> 2) & 7, m = x & 3;
  const double v = e == 0 ? std::ldexp(m / 4.0, 1 - 3) : std::ldexp(1.0 + m / 4.0, e - 3);
  return s ? -v : v;
}
uint32_t field6(const uint32_t w[6], int i) { // field i (0..31) of a 6-dword tuple
  const int bit = 6 * i, wd = bit >> 5, o = bit & 31;
  uint64_t v = w[wd] >> o;
  if (o > 26) v |= (uint64_t)w[wd + 1] << (32 - o);
  return (uint32_t)(v & 63u);
}

// The fp6 form: per polarization, the E operand tuples of pack_e_f6 read as
// E2M3 (times the A scale 2^3) against the weight tuple of pack_weights_f6
// read as E3M2 give the exact raw sums plus the same nibble-offset
// corrections; every one of the 16 dishes is used once. Also checks the
// codes themselves and that the fields' top bits stay clear.
void check_fp6_operands() {
  for (int n = 0; n < 16; ++n) expect(dec_e2m3((uint32_t)n) * 8.0 == n, "e2m3 of a nibble", n, 0);
  for (int w = -8; w <= 8; ++w) expect(dec_e3m2(e3m2_code(w)) == (double)w, "e3m2 of w", (long long)dec_e3m2(e3m2_code(w)), w);
  int seen[16] = {0};
  for (int j = 0; j < 16; ++j) ++seen[f6_dish(j)];
  for (int d = 0; d < 16; ++d) expect(seen[d] == 1, "fp6 dish used once", seen[d], 1);
  for (int trial = 0; trial < 40000; ++trial) {
    uint8_t e[32], a[16];
    for (int i = 0; i < 32; ++i) e[i] = (uint8_t)next();
    for (int i = 0; i < 16; ++i) a[i] = (uint8_t)next();
    if (trial < 256) {
      std::memset(e, trial, sizeof e);
      std::memset(a, trial & 1 ? 0x00 : 0xff, sizeof a);
    }
    uint32_t ew[8];
    std::memcpy(ew, e, 32);
    const v4i e0{(int)ew[0], (int)ew[1], (int)ew[2], (int)ew[3]}, e1{(int)ew[4], (int)ew[5], (int)ew[6], (int)ew[7]};
    uint32_t W[6];
    pack_weights_f6(a, a + 8, W);
    int sum_ar = 0, sum_ai = 0;
    for (int d = 0; d < 16; ++d) {
      sum_ar += bfo_dec_re(a[d]);
      sum_ai += bfo_dec_im(a[d]);
    }
    int corr_re, corr_im;
    corrections(sum_ar, sum_ai, &corr_re, &corr_im);
    for (int p = 0; p < 2; ++p) {
      v6i v2, v1;
      pack_e_f6(e0, e1, (uint32_t)p * 0x01010101u, v2, v1);
      int want_re = 0, want_im = 0;
      for (int d = 0; d < 16; ++d) {
        const int er = bfo_dec_re(e[2 * d + p]), ei = bfo_dec_im(e[2 * d + p]);
        const int ar = bfo_dec_re(a[d]), ai = bfo_dec_im(a[d]);
        want_re += ar * er - ai * ei;
        want_im += ar * ei + ai * er;
      }
      for (int variant = 0; variant < 2; ++variant) {
        const v6i op = variant ? v2 : v1;
        uint32_t ow[6];
        for (int i = 0; i < 6; ++i) ow[i] = (uint32_t)op[i];
        double f = 0;
        for (int i = 0; i < 32; ++i) {
          const uint32_t c = field6(ow, i);
          expect(c < 16, "fp6 E field top bits clear", c, 15);
          f += dec_e2m3(c) * 8.0 * dec_e3m2(field6(W, i));
        }
        const int got = (int)f - (variant ? corr_im : corr_re);
        const int want = variant ? want_im : want_re;
        expect(f == std::floor(f) && got == want, variant ? "fp6 raw im" : "fp6 raw re", got, want);
      }
    }
  }
}

// Gains the fast path takes: finite, |G raw| < 2^30 for |raw| <= 2^17.
void check_requant_gain(float g, int raw_lo, int raw_hi, int stride) {
  for (int raw = raw_lo; raw <= raw_hi; raw += stride) {
    const int want = bfo_requant(g, raw);
    const int exact = (int)requant_exact(g, raw);
    const int fast = (int)requant_fast(g, raw);
    const int bits = (int)(requant_bits(g, raw) & 0xffu) - 8;
    expect(exact == want, "requant_exact", exact, want);
    expect(fast == want, "requant_fast", fast, want);
    expect(bits == want, "requant_bits", bits, want);
  }
}

void check_requant() {
  float G[96];
  bfo_setup_gain(G, 96, 512); // the reference's gain formula (arraysizes.cxx:31-35)
  for (int b = 0; b < 96; b += 7) check_requant_gain(G[b], -65536, 65536, 1);
  for (int k = 0; k <= 16; ++k) { // powers of two: every half-integer tie
    check_requant_gain(std::ldexp(1.0f, -k), -65536, 65536, 1);
    check_requant_gain(-std::ldexp(1.0f, -k), -65536, 65536, 1);
  }
  check_requant_gain(0.0f, -65536, 65536, 1);
  check_requant_gain(-0.0f, -100, 100, 1);
  for (int i = 0; i < 200; ++i) {
    const float g = std::ldexp((float)(next() >> 40) / (float)(1 << 24), -(int)(next() % 20)) * (next() & 1 ? 1 : -1);
    check_requant_gain(g, -131072, 131072, 1 + (int)(next() % 13));
  }
  // the exact form on gains the fast path refuses (the generic kernel)
  const float odd[] = {std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                       std::numeric_limits<float>::quiet_NaN(), 1e30f, -1e30f, 3e9f, 65536.0f, 1e-45f};
  for (float g : odd)
    for (int raw = -300; raw <= 300; ++raw) {
      const int want = bfo_requant(g, raw);
      const int exact = (int)requant_exact(g, raw);
      expect(exact == want, "requant_exact (odd gain)", exact, want);
    }
}

void check_packing() {
  for (int trial = 0; trial < 20000; ++trial) {
    const float g = std::ldexp(1.0f, -(int)(next() % 12));
    int r[8];
    uint32_t bytes[4], want_word = 0;
    for (int i = 0; i < 8; ++i) r[i] = (int)(next() % 4001) - 2000;
    for (int i = 0; i < 4; ++i) {
      bytes[i] = jbyte_bits(requant_bits(g, r[2 * i]), requant_bits(g, r[2 * i + 1]));
      const uint8_t want = bfo_enc(bfo_requant(g, r[2 * i]), bfo_requant(g, r[2 * i + 1]));
      expect((bytes[i] & 0xffu) == want, "jbyte_bits", bytes[i] & 0xffu, want);
      want_word |= (uint32_t)want << (8 * i);
    }
    const uint32_t word = pack_low_bytes(bytes[0], bytes[1], bytes[2], bytes[3]);
    expect(word == want_word, "pack_low_bytes", word, want_word);
  }
}

} // namespace

int main() {
  check_codec();
  check_perm();
  check_operands();
  check_fp6_operands();
  check_requant();
  check_packing();
  if (failures) {
    std::printf("%d failures\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}

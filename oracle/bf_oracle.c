/*
 * bf_oracle.c -- CPU restatement of the reference form-beams path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline). The product
 * library never links or calls this file. See bf_oracle.h for the list of
 * reference citations and for what pins each function.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include "bf_oracle.h"

#include <math.h>
#include <omp.h>

/* ucomplex4 constructor: data = (((uchar)re << 4) | (im & 0xf)) ^ 0x88,
 * truncated to 8 bits (icomplex4.hxx:33). */
uint8_t bfo_enc(int re, int im) {
  unsigned v = ((((unsigned)(unsigned char)(signed char)re) << 4) | ((unsigned)im & 0x0fu)) ^ 0x88u;
  return (uint8_t)(v & 0xffu);
}

/* ucomplex4::real(): (signed char)(data ^ 0x88) >> 4 (icomplex4.hxx:36). */
int bfo_dec_re(uint8_t x) { return (int)((signed char)(uint8_t)(x ^ 0x88u)) >> 4; }

/* ucomplex4::imag(): (signed char)((uchar)(data ^ 0x88) << 4) >> 4 (icomplex4.hxx:37). */
int bfo_dec_im(uint8_t x) { return (int)((signed char)(uint8_t)((uint8_t)(x ^ 0x88u) << 4)) >> 4; }

/* Adler-32, one byte at a time, modulus 65521 (adler32.c:5-20). */
uint32_t bfo_adler32(const uint8_t *data, size_t len) {
  const uint32_t mod = 65521u;
  uint32_t a = 1, b = 0;
  for (size_t i = 0; i < len; ++i) {
    a = (a + data[i]) % mod;
    b = (b + a) % mod;
  }
  return (b << 16) | a;
}

/* G[n] = float(n) / Gsize / ndishes * (15 + n % 15) / 30, evaluated left to
 * right in float32 (arraysizes.cxx:31-32). */
void bfo_setup_gain(float *G, size_t nG, size_t ndishes) {
  for (size_t n = 0; n < nG; ++n) {
    float g = (float)n;
    g = g / (float)nG;
    g = g / (float)ndishes;
    g = g * (float)(15 + n % 15);
    g = g / (float)30;
    G[n] = g;
  }
}

/* E[n] = A[n] = ucomplex4(n % 15 - 7, (n + 1) % 15 - 7) over the byte index n
 * (arraysizes.cxx:16-30); the size_t wrap-around followed by the signed-char
 * conversion yields (n mod 15) - 7. */
void bfo_setup_pattern(uint8_t *E, size_t nE, uint8_t *A, size_t nA, float *G, size_t nG, size_t ndishes) {
  if (E) {
#pragma omp parallel for schedule(static)
    for (size_t n = 0; n < nE; ++n) E[n] = bfo_enc((int)(n % 15) - 7, (int)((n + 1) % 15) - 7);
  }
  if (A) {
    for (size_t n = 0; n < nA; ++n) A[n] = bfo_enc((int)(n % 15) - 7, (int)((n + 1) % 15) - 7);
  }
  if (G) bfo_setup_gain(G, nG, ndishes);
}

/* SplitMix64 (Steele, Lea, Flood 2014): output number `counter` (0-based) of
 * the generator seeded with `seed`. */
uint64_t bfo_splitmix64_at(uint64_t seed, uint64_t counter) {
  uint64_t z = seed + (counter + 1u) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Uniform in [-7, 7]: a 32-bit half h maps to ((h * 15) >> 32) - 7. Low half
 * gives the real part, high half the imaginary part. */
void bfo_fill_uniform(uint8_t *buf, size_t n, uint64_t seed, uint64_t offset) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) {
    const uint64_t z = bfo_splitmix64_at(seed, offset + i);
    const int re = (int)(((z & 0xffffffffull) * 15u) >> 32) - 7;
    const int im = (int)(((z >> 32) * 15u) >> 32) - 7;
    buf[i] = bfo_enc(re, im);
  }
}

void bfo_fill_phase_ramp(uint8_t *A, size_t F, size_t B, size_t D) {
  const double two_pi = 6.283185307179586476925286766559;
  for (size_t f = 0; f < F; ++f)
    for (size_t b = 0; b < B; ++b)
      for (size_t d = 0; d < D; ++d) {
        const double th = two_pi * (double)((b * d) % 97) / 97.0;
        const int re = (int)round(7.0 * cos(th));
        const int im = (int)round(7.0 * sin(th));
        A[(f * B + b) * D + d] = bfo_enc(re, im);
      }
}

/* max(-7, min(7, int(lrint(G * float(raw))))) (cpu.cxx:37-39). The float
 * product rounds to nearest, lrintf rounds ties to even, and the long -> int
 * conversion keeps the low 32 bits, exactly as the reference compiles on
 * x86-64. */
int bfo_requant(float gain, int raw) {
  const float y = gain * (float)raw;
  const int q = (int)lrintf(y);
  return q < -7 ? -7 : q > 7 ? 7 : q;
}

static inline size_t j_offset(int jlayout, size_t b, size_t f, size_t p, size_t t, size_t T, size_t F, size_t P) {
  /* Jlinear: ((b F + f) P + p) T + t (arraysizes.hxx:98);
   * J2linear: ((b F + f) T + t) P + p (arraysizes.hxx:109). */
  return jlayout == BFO_J_BFPT ? ((b * F + f) * P + p) * T + t : ((b * F + f) * T + t) * P + p;
}

/* The dish loop of cpu.cxx:24-35 for P == 2 (the reference's only
 * polarization count, arraysizes.hxx:25), with the stride known to the
 * compiler so that it vectorizes the loop as it does the reference's
 * compile-time shapes; cloned for AVX-512BW / AVX2 hosts. */
__attribute__((target_clones("avx512f", "avx2", "default"))) static void
dish_sum_p2(const uint8_t *Arow, const uint8_t *Erow, size_t D, int *out_re, int *out_im) {
  int raw_re = 0, raw_im = 0;
  for (size_t d = 0; d < D; ++d) {
    const uint8_t a = Arow[d];
    const uint8_t e = Erow[2 * d];
    const int are = (int)((signed char)(uint8_t)(a ^ 0x88u)) >> 4;
    const int aim = (int)((signed char)(uint8_t)((uint8_t)(a ^ 0x88u) << 4)) >> 4;
    const int ere = (int)((signed char)(uint8_t)(e ^ 0x88u)) >> 4;
    const int eim = (int)((signed char)(uint8_t)((uint8_t)(e ^ 0x88u) << 4)) >> 4;
    raw_re += are * ere - aim * eim;
    raw_im += are * eim + aim * ere;
  }
  *out_re = raw_re;
  *out_im = raw_im;
}

/* Any P. */
static void dish_sum(const uint8_t *Arow, const uint8_t *Erow, size_t D, size_t P, int *out_re, int *out_im) {
  int raw_re = 0, raw_im = 0;
  for (size_t d = 0; d < D; ++d) {
    const uint8_t a = Arow[d];
    const uint8_t e = Erow[d * P];
    const int are = bfo_dec_re(a), aim = bfo_dec_im(a);
    const int ere = bfo_dec_re(e), eim = bfo_dec_im(e);
    raw_re += are * ere - aim * eim;
    raw_im += are * eim + aim * ere;
  }
  *out_re = raw_re;
  *out_im = raw_im;
}

/* One output (cpu.cxx:24-41): A at Alinear, E at Elinear, G at Glinear. */
static inline uint8_t one_output(const uint8_t *E, const uint8_t *A, const float *G, size_t b, size_t f, size_t p,
                                 size_t t, size_t F, size_t D, size_t P, size_t B) {
  const uint8_t *Arow = A + (f * B + b) * D;              /* Alinear(f, b, d) = (f B + b) D + d */
  const uint8_t *Erow = E + ((t * F + f) * D) * P + p;    /* Elinear(t, f, d, p) = ((t F + f) D + d) P + p */
  int raw_re, raw_im;
  if (P == 2)
    dish_sum_p2(Arow, Erow, D, &raw_re, &raw_im);
  else
    dish_sum(Arow, Erow, D, P, &raw_re, &raw_im);
  const float g = G[f * B + b];                          /* Glinear(f, b) = b + B f */
  return bfo_enc(bfo_requant(g, raw_re), bfo_requant(g, raw_im));
}

int bfo_form_beams(uint8_t *J, const uint8_t *E, const uint8_t *A, const float *G, size_t T, size_t F, size_t D,
                   size_t P, size_t B, int jlayout, int nthreads) {
  if (jlayout != BFO_J_BFPT && jlayout != BFO_J_BFTP) return -1;
  if (T == 0 || F == 0 || D == 0 || P == 0 || B == 0) return 0;
  const long nb = (long)B;
  /* Parallel over beams, then f, p, t in that order (cpu.cxx:18-22). */
  if (nthreads > 0) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long b = 0; b < nb; ++b)
      for (size_t f = 0; f < F; ++f)
        for (size_t p = 0; p < P; ++p)
          for (size_t t = 0; t < T; ++t)
            J[j_offset(jlayout, (size_t)b, f, p, t, T, F, P)] = one_output(E, A, G, (size_t)b, f, p, t, F, D, P, B);
  } else {
#pragma omp parallel for schedule(static)
    for (long b = 0; b < nb; ++b)
      for (size_t f = 0; f < F; ++f)
        for (size_t p = 0; p < P; ++p)
          for (size_t t = 0; t < T; ++t)
            J[j_offset(jlayout, (size_t)b, f, p, t, T, F, P)] = one_output(E, A, G, (size_t)b, f, p, t, F, D, P, B);
  }
  return 0;
}

int bfo_form_beams_points(uint8_t *out, const uint64_t *idx, size_t npoints, const uint8_t *E, const uint8_t *A,
                          const float *G, size_t T, size_t F, size_t D, size_t P, size_t B, int jlayout) {
  if (jlayout != BFO_J_BFPT && jlayout != BFO_J_BFTP) return -1;
  const uint64_t total = (uint64_t)B * F * P * T;
  for (size_t i = 0; i < npoints; ++i)
    if (idx[i] >= total) return -1;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < npoints; ++i) {
    uint64_t r = idx[i];
    size_t b, f, p, t;
    if (jlayout == BFO_J_BFPT) {
      t = r % T; r /= T;
      p = r % P; r /= P;
      f = r % F; b = r / F;
    } else {
      p = r % P; r /= P;
      t = r % T; r /= T;
      f = r % F; b = r / F;
    }
    out[i] = one_output(E, A, G, b, f, p, t, F, D, P, B);
  }
  return 0;
}

/*
 * bf_oracle.h -- CPU restatement of the reference beamformer's form-beams path.
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE. Only tests/, the smoke check
 * in __graft_entry__.py and the cpu_baseline leg of bench.py load it, and only
 * as the checker (or, for cpu_baseline, as the timed CPU comparator). The
 * product path (libbeamform_amd.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates. Byte encodings,
 * index layouts and arithmetic follow the reference exactly:
 *   ucomplex4 codec ......... icomplex4.hxx:22-40
 *   layouts E/J/J2/A/G ...... arraysizes.hxx:74-142
 *   setup() pattern + gain .. arraysizes.cxx:8-36
 *   checksum flow ........... arraysizes.cxx:54-88, adler32.c:5-20
 *   form_beams .............. cpu.cxx:15-46
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   bfo_enc / bfo_dec ....... reference static_asserts (icomplex4.hxx:62-73),
 *                             all-256-byte table generated by the compiled
 *                             reference header (tests/golden/codec.json)
 *   bfo_adler32 ............. RFC 1950 known answers ("Wikipedia" = 0x11E60398),
 *                             the compiled reference adler32.c (golden fixture)
 *   bfo_setup_pattern ....... reference checksum table (arraysizes.cxx:54-59)
 *                             through bfo_form_beams
 *   bfo_form_beams .......... reference checksums at 6+ shapes, the compiled
 *                             reference cpu.cxx on seeded inputs
 *                             (tests/golden/ref_seeded.json), brute-force complex
 *                             arithmetic on tiny shapes, impulse closed forms,
 *                             ties-to-even and saturation special cases
 *   bfo_fill_uniform ........ defined here (no reference counterpart): pinned by
 *                             its own published test vectors and value-range /
 *                             uniformity checks; parity is by construction
 *                             because host and device consume the same bytes
 *   bfo_fill_phase_ramp ..... defined here from the workload description
 *                             (round(7cos), round(7sin) of 2*pi*(b*d mod 97)/97):
 *                             pinned by closed-form values at b*d mod 97 == 0
 *                             and symmetry checks
 */
#ifndef BF_ORACLE_H
#define BF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* J layouts: [b][f][p][t] (Jlinear, arraysizes.hxx:91-101) and
 * [b][f][t][p] (J2linear, arraysizes.hxx:103-112). */
enum { BFO_J_BFPT = 0, BFO_J_BFTP = 1 };

uint8_t bfo_enc(int re, int im);
int bfo_dec_re(uint8_t x);
int bfo_dec_im(uint8_t x);

uint32_t bfo_adler32(const uint8_t *data, size_t len);

/* The reference's deterministic inputs (arraysizes.cxx:16-32). */
void bfo_setup_pattern(uint8_t *E, size_t nE, uint8_t *A, size_t nA, float *G, size_t nG, size_t ndishes);
void bfo_setup_gain(float *G, size_t nG, size_t ndishes);

/* Seeded inputs: byte n = enc(re, im) with re, im uniform in [-7, 7] drawn
 * from the SplitMix64 output for counter (offset + n). */
void bfo_fill_uniform(uint8_t *buf, size_t n, uint64_t seed, uint64_t offset);
uint64_t bfo_splitmix64_at(uint64_t seed, uint64_t counter);

/* Phase-ramp weights A[f][b][d] = enc(round(7 cos th), round(7 sin th)),
 * th = 2 pi ((b*d) mod 97) / 97, identical for every f. */
void bfo_fill_phase_ramp(uint8_t *A, size_t F, size_t B, size_t D);

/* form_beams (cpu.cxx:15-46) with runtime shapes. nthreads <= 0 uses the
 * OpenMP default. Returns 0, or -1 for an invalid shape / layout. */
int bfo_form_beams(uint8_t *J, const uint8_t *E, const uint8_t *A, const float *G, size_t T, size_t F, size_t D,
                   size_t P, size_t B, int jlayout, int nthreads);

/* Same arithmetic for selected outputs only: J byte offsets idx[i] (in the
 * given layout) are recomputed into out[i]. Used for spot checks of large
 * outputs. */
int bfo_form_beams_points(uint8_t *out, const uint64_t *idx, size_t npoints, const uint8_t *E, const uint8_t *A,
                          const float *G, size_t T, size_t F, size_t D, size_t P, size_t B, int jlayout);

/* Requantisation of one raw sum (cpu.cxx:37-39). */
int bfo_requant(float gain, int raw);

#ifdef __cplusplus
}
#endif

#endif

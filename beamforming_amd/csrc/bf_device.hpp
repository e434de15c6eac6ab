// bf_device.hpp -- per-lane arithmetic of the gfx950 kernels. Every function
// also compiles for the host, where tests/device_math_check.cpp checks it
// against the oracle.
//
// Everything here is integer/byte work except the requantisation, which
// reproduces the reference's float semantics bit for bit (cpu.cxx:37-39).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#define BF_HD __host__ __device__ __forceinline__

namespace bf {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// ucomplex4 codec (icomplex4.hxx:33-37): offset binary, real part in the high
// nibble, each part stored as value + 8.
BF_HD int dec_re(uint32_t x) { return (int)((int8_t)(uint8_t)(x ^ 0x88u)) >> 4; }
BF_HD int dec_im(uint32_t x) { return (int)((int8_t)(uint8_t)((x ^ 0x88u) << 4)) >> 4; }
BF_HD uint32_t enc(int re, int im) { return ((((uint32_t)re & 0xfu) << 4) | ((uint32_t)im & 0xfu)) ^ 0x88u; }

// ---------------------------------------------------------------------------
// GEMM formulation.
//
// For one frequency, with E-rows = time samples and columns = beams:
//   raw_re[t,b] = sum_d Ar Er - Ai Ei,   raw_im[t,b] = sum_d Ar Ei + Ai Er.
// The weights are held once as int8 pairs W = (Ar, Ai) per dish. The E side
// is fed to the int8 MFMA as the raw nibbles s = value + 8 (0..15), in two
// variants that share W:
//   v1 = (s_re, 15 - s_im)  ->  W.v1 = raw_re + 8 sum Ar + 7 sum Ai
//   v2 = (s_im, s_re)       ->  W.v2 = raw_im + 8 sum Ar + 8 sum Ai
// The constant offsets depend only on (f, b); the accumulators start at their
// negation, so they end at exactly raw_re / raw_im (exact int32 arithmetic).
//
// MFMA v_mfma_i32_16x16x64_i8 operands: lane l holds 16 int8 of row l % 16
// (E: time sample) or column l % 16 (W: beam), k-values 16 (l / 16) ..
// + 15 of the 64. Per lane and MFMA the 16 k-values cover 8 dishes as four
// dwords: [re d0-3 | re d4-7 | im d0-3 | im d4-7] on the weight side. On the E side the nibble dwords
// u = [lo_a lo_b hi_a hi_b nlo_a nlo_b] (a = dishes 0-3, b = 4-7) give both
// variants as overlapping register windows: v2 = u[0..3], v1 = u[2..5].

// Byte permute: result byte i = byte sel_i of the 8-byte value (hi:lo),
// selectors 0-3 address lo, 4-7 address hi, 12 gives 0x00 (v_perm_b32).
// Other selector values are not used here.
BF_HD uint32_t perm_bytes(uint32_t hi, uint32_t lo, uint32_t sel) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_perm(hi, lo, sel);
#else
  const uint64_t v = ((uint64_t)hi << 32) | lo;
  uint32_t r = 0;
  for (int i = 0; i < 4; ++i) {
    const uint32_t s = (sel >> (8 * i)) & 0xffu;
    const uint32_t byte = s < 8 ? (uint32_t)((v >> (8 * s)) & 0xffu) : 0u;
    r |= byte << (8 * i);
  }
  return r;
#endif
}

typedef int v8i __attribute__((ext_vector_type(8)));

// Keeps one polarization of 16 packed E bytes (8 dishes x 2 polarizations:
// byte 2k+p = dish k, polarization p; sel = 0x06040200 keeps p = 0,
// 0x07050301 keeps p = 1) and splits its nibbles into the operand window
// u = [lo_a lo_b hi_a hi_b nlo_a nlo_b . .]: v2 = u[0..3], v1 = u[2..5].
BF_HD v8i unpack_e_pol(const v4i e, uint32_t sel) {
  const uint32_t qa = perm_bytes((uint32_t)e.y, (uint32_t)e.x, sel); // dishes 0-3
  const uint32_t qb = perm_bytes((uint32_t)e.w, (uint32_t)e.z, sel); // dishes 4-7
  const uint32_t lo_a = qa & 0x0f0f0f0fu, hi_a = (qa >> 4) & 0x0f0f0f0fu;
  const uint32_t lo_b = qb & 0x0f0f0f0fu, hi_b = (qb >> 4) & 0x0f0f0f0fu;
  return v8i{(int)lo_a, (int)lo_b, (int)hi_a, (int)hi_b, (int)(lo_a ^ 0x0f0f0f0fu), (int)(lo_b ^ 0x0f0f0f0fu), 0, 0};
}
BF_HD v4i op_v1(const v8i u) { return __builtin_shufflevector(u, u, 2, 3, 4, 5); }
BF_HD v4i op_v2(const v8i u) { return __builtin_shufflevector(u, u, 0, 1, 2, 3); }

// K steps of 64 dishes on v_mfma_scale_f32_16x16x128_f8f6f4 with fp6 operands
// (below) for KS = 4 .. 32; the int8 MFMA for KS = 2 and 64 (at KS = 64 the
// longer unrolled loop of the 16x16x128 form does not fit the registers).
BF_HD constexpr bool mfma_f6(int ks) { return ks >= 4 && ks <= 32; }
// The MFMA instruction of a K-step count, as plan and batch descriptions name it.
inline const char *mfma_name(int ks) {
  return mfma_f6(ks) ? "mfma_f32_16x16x128_f8f6f4(e2m3 x e3m2)" : "mfma_i32_16x16x64_i8";
}

// Weight operand for one lane: 8 packed A bytes (dishes d0..d0+7, zero-padded
// by the caller) -> [Ar d0-3 | Ar d4-7 | Ai d0-3 | Ai d4-7] as int8.
BF_HD v4i pack_weights(const uint8_t a[8]) {
  uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t x = a[4 * q + i];
      w[q + 0] |= ((uint32_t)dec_re(x) & 0xffu) << (8 * i);
      w[q + 2] |= ((uint32_t)dec_im(x) & 0xffu) << (8 * i);
    }
  return v4i{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
}

// ---------------------------------------------------------------------------
// fp6 form (v_mfma_scale_f32_16x16x128_f8f6f4 with A in E2M3, B in E3M2).
// In a register-only loop on all CUs (random operands in these ranges) the
// instruction runs 7.57 POP/s at 1341 W with fp6 operands against 4.82 POP/s
// at 1285 W with fp8 operands: two thirds of the energy per operation, and
// the launch runs at the board's power limit. A lane's operand is 32 six-bit
// fields, field i at bits 6 i of a 6-dword tuple (k = 32 (lane / 16) + i):
// 16 dishes x (re, im) of a 64-dish step (dishes 8 kg .. 8 kg + 7 of each of
// its two 32-dish halves).
//  * E: a nibble n = 0..15 in the low 4 bits of a field (top 2 bits zero) is
//    the E2M3 code of n / 8 (0..7 subnormal, 8..15 exponent 1), exactly; the
//    MFMA's A scale 2^3 makes every product the integer n w.
//  * Weights: w = -8..7 as E3M2 codes (sign, 3 exponent bits of bias 3, 2
//    mantissa bits: every such w is exact).
// Every partial sum is an integer below 2^17: the fp32 accumulators end at
// the same raw sums as the int8 form's int32 accumulators, exactly.
//
// The operand halves. For one polarization the 16 dishes' nibbles fill two
// 16-field halves H1 (dwords 0..2) and H2 (dwords 3..5): field j of H1 holds
// one nibble of dish kF6Dish[j] and field j of H2 the other. Where H1 holds
// s_im ("case A") the weight halves are (Ar, Ai), where it holds s_re ("case
// B") they are (Ai, Ar); then
//   [H1, H2] . W                     = raw_im + 8 sum Ar + 8 sum Ai
//   [H2 ^ MR, H1 ^ MI] . W           = raw_re + 8 sum Ar + 7 sum Ai
// where the masks complement s_im (to 15 - s_im) in the case-B fields of H2
// (MR) and the case-A fields of H1 (MI): the same nibble-offset corrections
// as the other forms. The dish of field j (0 .. 7: dishes 8 kg + 0 .. 7 of
// the step's first 32 dishes, 8 .. 15: of its second 32) and its case:
BF_HD constexpr int f6_dish(int j) {
  constexpr int d[16] = {0, 3, 1, 4, 2, 5, 8, 6, 9, 7, 10, 13, 11, 14, 12, 15};
  return d[j];
}
BF_HD constexpr bool f6_case_b(int j) { return (0x64CCu >> j) & 1u; } // fields 2 3 6 7 10 13 14
// E3M2 code of an integer w in -8..8.
BF_HD uint32_t e3m2_code(int w) {
  const int a = w < 0 ? -w : w;
  int c = 0;
  if (a == 1) c = 12;       // 1.00 x 2^0
  else if (a == 2) c = 16;  // 1.00 x 2^1
  else if (a == 3) c = 18;  // 1.50 x 2^1
  else if (a >= 4) c = 20 + (a - 4); // 4..7: 1.00 .. 1.75 x 2^2; 8: 1.00 x 2^3
  return (uint32_t)(c | (w < 0 ? 32 : 0));
}
// Weight operand halves of one lane and 64-dish step (a0: the 8 dishes of the
// step's first 32, a1: of its second 32, ucomplex4 bytes, 0x88 where zero):
// WA (dwords 0..2) and WB (dwords 3..5), field j at bits 6 j.
BF_HD void pack_weights_f6(const uint8_t a0[8], const uint8_t a1[8], uint32_t w[6]) {
  for (int i = 0; i < 6; ++i) w[i] = 0;
  for (int j = 0; j < 16; ++j) {
    const int dd = f6_dish(j);
    const uint32_t x = dd < 8 ? a0[dd] : a1[dd - 8];
    const int ar = dec_re(x), ai = dec_im(x);
    const uint32_t ca = e3m2_code(f6_case_b(j) ? ai : ar), cb = e3m2_code(f6_case_b(j) ? ar : ai);
    const int bit = 6 * j, wd = bit >> 5, o = bit & 31;
    w[wd] |= ca << o;
    w[3 + wd] |= cb << o;
    if (o > 26) {
      w[wd + 1] |= ca >> (32 - o);
      w[3 + wd + 1] |= cb >> (32 - o);
    }
  }
}
// E operands of one polarization (sel_add = 0 or 1) of a 64-dish step from
// the lane's two 16-byte loads (r0..r3 = e0: dishes 8 kg + 0 .. 7 of the
// first 32, both polarizations, two dishes per dword; r4..r7 = e1). Each
// 32-bit piece of a half is two byte permutes of the raw dwords (the
// polarization's bytes, a dish placed twice where the two halves need its
// two nibbles at different shifts) merged under field masks:
//   H1: d0 = Pa & Ma | Pb << 6 & Mb,  d1 = Aq & MA | Bq >> 2 & MB,  d2 = Cq & MC | Dq << 2 & MD
//   H2: d0 = Pa >> 4 & Ma | Pb << 2 & Mb, d1 = Aq << 4 & MA | Bq >> 6 & MB, d2 = Cq << 4 & MC | Dq >> 2 & MD
// (field 5 straddles d0 and d1 of each half). v2 = [H1, H2], v1 = [H2 ^ MR,
// H1 ^ MI].
typedef int v6i __attribute__((ext_vector_type(6)));
// a | (b & m) in one v_bitop3_b32 (truth table 0xF8: inputs a, b, m as bits
// 2, 1, 0 of the row index); the compiler otherwise spends two ANDs and an OR.
BF_HD uint32_t or_masked(uint32_t a, uint32_t b, uint32_t m) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, m, 0xF8);
#else
  return a | (b & m);
#endif
}
BF_HD void pack_e_f6(const v4i e0, const v4i e1, uint32_t sel_add, v6i &v2, v6i &v1) {
  const uint32_t r0 = (uint32_t)e0.x, r1 = (uint32_t)e0.y, r2 = (uint32_t)e0.z, r3 = (uint32_t)e0.w;
  const uint32_t r4 = (uint32_t)e1.x, r5 = (uint32_t)e1.y, r6 = (uint32_t)e1.z, r7 = (uint32_t)e1.w;
  // byte selectors for polarization 0; polarization 1 is one byte up
  const uint32_t sa = 0x04020200u + sel_add, sb = 0x06040402u + sel_add, sc = 0x04040202u + sel_add;
  const uint32_t pa = perm_bytes(r1, r0, sa); // dishes 0 1 1 2
  const uint32_t pb = perm_bytes(r2, r1, sb); // dishes 3 4 4 5
  const uint32_t bq = perm_bytes(r3, r2, sb); // dishes 5 6 6 7
  const uint32_t aq = perm_bytes(r5, r4, sa); // dishes 8 9 9 10
  const uint32_t cq = perm_bytes(r6, r5, sc); // dishes 11 11 12 12
  const uint32_t dq = perm_bytes(r7, r6, sb); // dishes 13 14 14 15
  const uint32_t h10 = or_masked(pa & 0x0F00F00Fu, pb << 6, 0xC03C03C0u);
  const uint32_t h11 = or_masked(aq & 0xF00F00F0u, bq >> 2, 0x03C03C03u);
  const uint32_t h12 = or_masked(cq & 0x00F00F00u, dq << 2, 0x3C03C03Cu);
  const uint32_t h20 = or_masked((pa >> 4) & 0x0F00F00Fu, pb << 2, 0xC03C03C0u);
  const uint32_t h21 = or_masked((aq << 4) & 0xF00F00F0u, bq >> 6, 0x03C03C03u);
  const uint32_t h22 = or_masked((cq << 4) & 0x00F00F00u, dq >> 2, 0x3C03C03Cu);
  v2 = v6i{(int)h10, (int)h11, (int)h12, (int)h20, (int)h21, (int)h22};
  v1 = v6i{(int)(h20 ^ 0x003CF000u), (int)(h21 ^ 0xF0003CF0u), (int)(h22 ^ 0x00F3C000u),
           (int)(h10 ^ 0xCF0003CFu), (int)(h11 ^ 0x03CF0003u), (int)(h12 ^ 0x3C000F3Cu)};
}

// ---------------------------------------------------------------------------
// Requantisation, cpu.cxx:37-39: max(-7, min(7, int(lrint(G * float(raw))))).
//
// Fast form, valid when G is finite and |G * raw| < 2^31 for every raw the
// shape can produce (the plan checks this): clamping before rounding gives the
// same value because rint is monotonic and fixes +-7.
BF_HD float requant_fast(float g, int raw) {
  const float y = g * (float)raw;
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_rintf(__builtin_amdgcn_fmed3f(y, -7.0f, 7.0f));
#else
  return rintf(fminf(fmaxf(y, -7.0f), 7.0f));
#endif
}

// Exact form for any G: lrintf on x86-64 yields INT64_MIN for NaN and
// |y| >= 2^63, and the long -> int conversion keeps the low 32 bits.
BF_HD float requant_exact(float g, int raw) {
  const float y = g * (float)raw;
  const float ay = fabsf(y);
  int q;
  if (ay < 2147483648.0f)
    q = (int)rintf(y);
  else if (!(ay < 9223372036854775808.0f))
    q = 0;
  else
    q = (int)(uint32_t)(uint64_t)(int64_t)y; // y is integral here
  q = q < -7 ? -7 : (q > 7 ? 7 : q);
  return (float)q;
}

// Fast requantisation with the result in integer bits: the low byte of the
// returned word is requant_fast(g, raw) + 8, a nibble 1..15 (the ucomplex4
// offset-binary digit). c = clamp(g * raw, -7, 7) is added to
// M = 1.5 * 2^23 + 8: floats in [2^23, 2^24) are the integers, so the sum is
// rounded to the nearest integer, ties to even; M is even, so the tie goes
// to the even rint(c), as in rint. The float bits of M + 8 + rint(c) are
// bits(1.5 * 2^23) + 8 + rint(c), whose low 22 bits are 8 + rint(c).
BF_HD uint32_t requant_bits_f(float g, float raw) {
  const float y = g * raw;
#if defined(__HIP_DEVICE_COMPILE__)
  const float c = __builtin_amdgcn_fmed3f(y, -7.0f, 7.0f);
#else
  const float c = fminf(fmaxf(y, -7.0f), 7.0f);
#endif
  return __builtin_bit_cast(uint32_t, c + 12582920.0f);
}
// The same from an int32 raw sum (|raw| < 2^24, so the conversion is exact),
// and from a raw sum already held exactly as a float.
BF_HD uint32_t requant_bits(float g, int raw) { return requant_bits_f(g, (float)raw); }
BF_HD uint32_t requant_bits(float g, float raw) { return requant_bits_f(g, raw); }

// ucomplex4 byte of requant_bits pairs in the low byte: (re + 8) << 4 |
// (im + 8), i.e. enc(qr, qi) (icomplex4.hxx:33). Higher bytes are garbage.
BF_HD uint32_t jbyte_bits(uint32_t re_bits, uint32_t im_bits) { return (re_bits << 4) | im_bits; }

// Word of the low bytes of b0..b3 (byte i = low byte of bi).
BF_HD uint32_t pack_low_bytes(uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3) {
  const uint32_t t01 = perm_bytes(b1, b0, 0x0c0c0400u);
  const uint32_t t23 = perm_bytes(b3, b2, 0x0c0c0400u);
  return perm_bytes(t23, t01, 0x05040100u);
}

// Offsets that the E-side nibble encoding adds to the two accumulators.
BF_HD void corrections(int sum_ar, int sum_ai, int *corr_re, int *corr_im) {
  *corr_re = 8 * sum_ar + 7 * sum_ai;
  *corr_im = 8 * sum_ar + 8 * sum_ai;
}

// ---------------------------------------------------------------------------
// Work decomposition of the MFMA kernel.
constexpr int kWaves = 8;            // waves per workgroup (2 per SIMD)
constexpr int kThreads = 64 * kWaves;
constexpr int kTileT = 16;           // time samples per wave tile (MFMA M; one MFMA row set per polarization)
constexpr int64_t kMaxTimes = int64_t(1) << 34; // longest chunk: time tile indices stay below 2^30
constexpr int kDishesPerStep = 16;   // dishes per K step; one MFMA (K = 64 int8) takes two steps
constexpr int kBeamTile = 32;        // beams per beam tile (two 16-beam MFMA column groups)
constexpr int kMaxBeamTiles = 3;     // beam tiles per workgroup (NT), registers permitting

struct MfmaParams {
  const uint8_t *E; // [T][F][D][2]
  uint8_t *J;       // [B][F][2][T] or [B][F][T][2]
  const v4i *W;     // [F][KS / 2][2 NTP][64] weight operands (32-dish steps, 16-beam groups)
  const int *corr;  // [F][BP][2], BP = 32 NTP
  const float *G;   // [F][BP]
  int64_t T;
  int32_t F, D, B;
  int32_t NTP;       // 32-beam tiles over all beam groups (nbg * NT)
  int32_t nbg;       // beam groups (one workgroup column each)
  int32_t nchunks;   // time chunks per (f, beam group)
  int32_t tiles_per_chunk;
  int32_t nwork;     // F * nbg * nchunks
  int32_t per_xcd;   // ceil(nwork / 8)
  int32_t jlayout;   // 0 = [b][f][p][t], 1 = [b][f][t][p]
  int32_t wrow0;     // weight-table row of frequency 0 (a channel, for one-channel calls)
  int32_t stage_j;   // 1: J rows go through the LDS staging area (allocated in the launch's LDS)
};

// ---------------------------------------------------------------------------
// Address arithmetic of the tile pipeline, shared with the host bounds check
// (tests/bounds_check.cpp), which runs these same expressions for every wave
// and lane over random shapes and chunk lists.
//
// E: lane (row, piece) of a tile loads 16 bytes of time row `row` (0..15),
// bytes 16 piece .. + 15 of the row's 64-byte window of K-step pair j
// (dishes 32 j .. 32 j + 31, both polarizations). Rows past T re-read row
// T - 1 (their outputs are never stored).
BF_HD int64_t e_row_index(int64_t tile, int row, int64_t T) {
  const int64_t t = tile * kTileT + row;
  return t < T ? t : T - 1;
}
// Byte offset within the row of the piece's load, relative to byte
// 16 piece. With a dish tail (dt), pieces wholly past D (D is a multiple of
// 8, so a piece is either wholly inside or wholly past) re-read the row's
// last 16 bytes instead (their weights are zero): the offset is clamped to
// e_col_limit, the last offset whose 16 bytes lie inside the row.
BF_HD int e_col_limit(int piece, int D) { return 2 * D - 16 - 16 * piece; }
BF_HD int e_col_offset_lim(int j, int lim) { return 64 * j < lim ? 64 * j : lim; }
BF_HD int64_t e_col_offset(int j, int piece, int D, bool dt) {
  if (dt) return e_col_offset_lim(j, e_col_limit(piece, D));
  return 64 * j;
}
// First byte of J row (beam b, frequency f, polarization p) of
// J[b][F][p][T] ([b][f][p][t]) and of beam row (b, f) of J[b][F][T][2]
// ([b][f][t][p]).
BF_HD int64_t j_row_bfpt(int b, int F, int f, int p, int64_t T) { return (((int64_t)b * F + f) * 2 + p) * T; }
BF_HD int64_t j_row_bftp(int b, int F, int f, int64_t T) { return ((int64_t)b * F + f) * T * 2; }
// LDS staging area of J[b][f][p][t]: two slots (even and odd rounds) of
// 2 NT x 32 rows x 128 bytes, rows in (beam, polarization) order, the 16-byte
// pieces of a row rotated by its beam (which spreads the epilogue's writes of
// 16 beams over the LDS banks). stage_word: the word of lane kg of `wave`
// (times 16 wave + 4 kg .. + 3 of the round) in row srow; stage_piece_byte:
// the byte of piece seg (times 16 seg .. + 15 of the round) of row srow.
BF_HD int stage_word(int nt, int slot, int srow, int wave, int kg) {
  return slot * nt * 64 * 32 + srow * 32 + (wave ^ ((srow >> 1) & 7)) * 4 + kg;
}
BF_HD int stage_piece_byte(int nt, int slot, int srow, int seg) {
  return 16 * (slot * nt * 64 * 8 + srow * 8 + (seg ^ ((srow >> 1) & 7)));
}

// LDS staging area of J[b][f][t][p]: two slots of NT x 32 beam rows x 256
// bytes (8 tiles x 16 times x 2 polarizations), the 16-byte pieces of a row
// rotated by its beam. stage_pair_bftp: the byte of lane kg's 8 bytes (times
// 4 kg .. 4 kg + 3 of `wave`'s tile, polarizations interleaved) in row srow;
// stage_piece_bftp: the byte of piece seg (times 8 seg .. 8 seg + 7).
BF_HD int stage_pair_bftp(int nt, int slot, int srow, int wave, int kg) {
  return 8 * (slot * nt * 64 * 16 + srow * 32 + ((2 * wave + (kg >> 1)) ^ (srow & 15)) * 2 + (kg & 1));
}
BF_HD int stage_piece_bftp(int nt, int slot, int srow, int seg) {
  return 16 * (slot * nt * 64 * 8 + srow * 16 + (seg ^ (srow & 15)));
}

// Tiles [begin, end) of time chunk `chunk` of a plain launch.
BF_HD void chunk_tiles(int chunk, int tiles_per_chunk, int64_t ntiles, int64_t *begin, int64_t *end) {
  *begin = (int64_t)chunk * tiles_per_chunk;
  *end = ntiles < *begin + tiles_per_chunk ? ntiles : *begin + tiles_per_chunk;
}

// Work index of a workgroup. Workgroups are dispatched round-robin over the
// 8 XCDs; this gives each XCD a contiguous range of work items, so the beam
// groups of one time chunk (which read the same E rows) and consecutive
// chunks of one frequency (same weights) share that XCD's L2.
BF_HD int work_of_block(int block, int per_xcd) { return (block & 7) * per_xcd + (block >> 3); }

struct WorkItem {
  int f, bg, chunk;
};
BF_HD WorkItem decode_work(int work, int nbg, int nchunks) {
  WorkItem w;
  w.bg = work % nbg;
  const int rest = work / nbg;
  w.chunk = rest % nchunks;
  w.f = rest / nchunks;
  return w;
}

} // namespace bf

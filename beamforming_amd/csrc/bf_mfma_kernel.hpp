// bf_mfma_kernel.hpp -- the gfx950 form-beams kernel (MFMA, exact).
//
// GEMM per frequency: rows = time samples of one polarization (16 per MFMA
// block), columns = beams (16 per MFMA block), K = dishes x (re, im). The
// MFMA is v_mfma_scale_f32_16x16x128_f8f6f4 (64 dishes per instruction, fp6
// operands holding the int4 values exactly, fp32 accumulators holding exact
// integer sums) for 64 .. 512 dishes, and v_mfma_i32_16x16x64_i8 (32 dishes
// per instruction) for fewer or more: the launch runs at the board's power
// limit, where the fp6 form spends the least energy per operation.
//
// One workgroup = one (frequency, beam group of NT x 32 beams, time chunk):
//  * the prepared weight operands of the beam group (KS x NT KiB) are staged
//    once into LDS and shared by the 8 waves (2 per SIMD);
//  * each wave owns 16-sample time tiles. Per K step it loads the tile's
//    16 rows x 64 bytes of E (dishes 32 s .. 32 s + 31, both polarizations)
//    into a register ring running RP steps ahead, each lane (kg, col) =
//    (lane / 16, lane % 16) loading the 16 bytes of its own operand: row
//    (time) col, dishes 8 kg .. 8 kg + 7 (the four lanes of a row read its
//    64-byte segment); byte permutes and masked shifts turn the nibbles of
//    each polarization into the two operand variants (see bf_device.hpp);
//  * per K step (fp6: pair of steps) and 16-beam group: one LDS read of the
//    weight operand (fp6: two), shared by both polarizations, and per
//    polarization two MFMAs (real and imaginary accumulator);
//  * the accumulators (int32, or fp32 holding integers exactly) start at
//    minus the per-beam nibble offsets, so
//    they end at the exact complex sums of cpu.cxx:24-35; the fused epilogue
//    applies the float gain, the +-7 clamp and round-to-nearest-even
//    (cpu.cxx:37-39) and packs ucomplex4 bytes: each accumulator holds 4
//    consecutive times of one (beam, polarization), one 4-byte J word;
//  * the 8 waves advance in rounds of 8 consecutive tiles. When the launch
//    stages J (room in LDS, and the configuration asks for it), each wave
//    writes its words into a double-buffered LDS staging area and, after
//    the round's barrier, the workgroup stores every row's bytes of the
//    round (8 tiles: 128 bytes per (beam, polarization) row of [b][f][p][t],
//    256 bytes per beam row of [b][f][t][p]) with whole-line writes;
//    otherwise each lane stores its words directly. The J layout is a
//    template parameter (JL), so each kernel holds one epilogue.
// Rows past T and dishes past D are clamped, always-in-bounds loads: the
// outputs of rows past T are never stored and the weights of dishes past D
// are zero. The plan routes gains outside the fast requantisation range
// (non-finite, or |G * raw| >= 2^30 possible) to the generic kernel.
#pragma once

#include "bf_device.hpp"
#include "bf_items.hpp"

#include <type_traits>

namespace bf {

typedef int v16i __attribute__((ext_vector_type(16)));

// Global-memory views. E and J are always device memory, but pointers read
// from a chunk table are generic to the compiler; generic (flat) accesses
// also count as LDS operations, and a pending one makes every wait a full
// drain of both counters. Every E load and J store goes through these.
#define BF_GLOBAL __attribute__((address_space(1)))
template <class T> __device__ __forceinline__ const BF_GLOBAL T *gload_ptr(const void *p) {
  return (const BF_GLOBAL T *)p;
}
template <class T> __device__ __forceinline__ BF_GLOBAL T *gstore_ptr(void *p) { return (BF_GLOBAL T *)p; }

// Vector types without an alignment promise: gfx950 runs in unaligned
// access mode, so a 16- or 8-byte load or store at any address is one
// instruction (split by the memory system where it crosses a line). E loads
// and J stores use them, so E and J may sit at any byte address; the code is
// the same as for aligned types.
typedef v4i v4i_u __attribute__((aligned(1)));
typedef int v2i __attribute__((ext_vector_type(2)));
typedef v2i v2i_u __attribute__((aligned(1)));
typedef uint32_t u32_u __attribute__((aligned(1)));

// Stores 16 bytes (w0..w3, little-endian) at dst, keeping only bytes i < n.
__device__ __forceinline__ void store_row16(uint8_t *dst, uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, int n) {
  if (n >= 16) {
    *gstore_ptr<v4i_u>(dst) = v4i{(int)w0, (int)w1, (int)w2, (int)w3};
  } else {
    const uint32_t w[4] = {w0, w1, w2, w3};
    BF_GLOBAL uint8_t *d = gstore_ptr<uint8_t>(dst);
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if (i < n) d[i] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
  }
}

// Stores 8 bytes (w0, w1) at dst, keeping bytes i < n.
__device__ __forceinline__ void store_row8(uint8_t *dst, uint32_t w0, uint32_t w1, int n) {
  if (n >= 8) {
    *gstore_ptr<v2i_u>(dst) = v2i{(int)w0, (int)w1};
  } else {
    const uint32_t w[2] = {w0, w1};
    BF_GLOBAL uint8_t *d = gstore_ptr<uint8_t>(dst);
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (i < n) d[i] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
  }
}

// LDS of a workgroup: [KS][NT][64] weight operands, the beam-column nibble
// offsets and gains of the beam group, then, when the launch stages J (and it fits in the 160 KiB of a CU), the J
// staging area: 2 rounds x (NT x 32 beams x 2 polarizations) rows x 128 bytes
// ([b][f][p][t]), or 2 rounds x NT x 32 beam rows x 256 bytes ([b][f][t][p]).
constexpr size_t kLdsPerCu = 160 * 1024;
BF_HD constexpr size_t mfma_lds_unstaged(int nt, int ks) {
  return (size_t)ks * nt * 64 * sizeof(v4i) + (size_t)nt * 32 * (sizeof(int2) + sizeof(float));
}
BF_HD constexpr size_t mfma_lds_staging(int nt) { return (size_t)2 * nt * 64 * 128; }
BF_HD constexpr bool mfma_stages_j(int nt, int ks) {
  return mfma_lds_unstaged(nt, ks) + mfma_lds_staging(nt) <= kLdsPerCu;
}
BF_HD constexpr size_t mfma_lds_bytes(int nt, int ks) {
  return mfma_lds_unstaged(nt, ks) + (mfma_stages_j(nt, ks) ? mfma_lds_staging(nt) : 0);
}

// Stages row `wrow` (a frequency, or a channel of a channel table) of the
// prepared weights, beam group bg, into LDS. All threads of the workgroup
// take part; the caller synchronises.
template <int NT, int KS>
__device__ __forceinline__ void stage_weights(const v4i *W, const int *corr, const float *G, int NTP, int wrow, int bg,
                                              v4i *Wl, int2 *Cl, float *Gl) {
  // W [row][KS / 2][2 NTP][64]: per 32-dish step, the 16-beam groups of all
  // beam tiles; the workgroup's 2 NT groups of every step go to LDS as
  // Wl [KS / 2][2 NT][64].
  constexpr int KQ = KS / 2, NQ = 2 * NT;
  const int NQP = 2 * NTP;
  const v4i *src = W + ((size_t)wrow * KQ * NQP + (size_t)bg * NQ) * 64;
  // Every load is issued before the first LDS write, so the copy costs one
  // memory latency instead of one per 8 KiB (up to 16 v4i per thread).
  static_assert(NT * 32 <= kThreads, "one offset and gain per thread");
  const size_t g0 = (size_t)wrow * NTP * 32 + (size_t)bg * NT * 32;
  const bool cg = threadIdx.x < NT * 32;
  int2 cv = make_int2(0, 0);
  float gv = 0.0f;
  if (cg) {
    cv = make_int2(corr[(g0 + threadIdx.x) * 2 + 0], corr[(g0 + threadIdx.x) * 2 + 1]);
    gv = G[g0 + threadIdx.x];
  }
  constexpr int N = KQ * NQ * 64;
  constexpr int PER = (N + kThreads - 1) / kThreads;
  // The thread index made opaque, so that the PER source offsets are
  // computed here: hoisted out of a caller's loop (the batch kernel stages
  // weights per segment) they were held in registers, spilled, and reloaded
  // between the loads with a full wait each, which turned one memory latency
  // into six (10 us instead of 2.5 at D = 512, NT = 3).
  int tid = threadIdx.x;
  __asm__ volatile("" : "+v"(tid));
  v4i tmp[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = tid + k * kThreads;
    if (N % kThreads == 0 || i < N) {
      const int s = i / (NQ * 64);
      const int rem = i - s * (NQ * 64);
      tmp[k] = src[(size_t)s * NQP * 64 + rem];
    }
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = tid + k * kThreads;
    if (N % kThreads == 0 || i < N) Wl[i] = tmp[k];
  }
  if (cg) {
    // fp6 accumulators are floats: their start values, minus the offsets
    // (exact: |offset| < 2^24), are stored ready to use
    if constexpr (mfma_f6(KS))
      Cl[threadIdx.x] = make_int2(__float_as_int((float)-cv.x), __float_as_int((float)-cv.y));
    else
      Cl[threadIdx.x] = cv;
    Gl[threadIdx.x] = gv;
  }
}

// One stream of E rows (one frequency of one chunk) and where its J goes.
struct TileJob {
  const uint8_t *E;  // E byte of (t = 0, d = 0, p = 0) of this frequency
  int64_t e_tstride; // bytes between consecutive times (F D 2)
  uint8_t *J;        // J of the whole [B][F][..] array
  int64_t T;         // time samples of the stream
  int32_t F, f;      // frequencies in J, and this stream's frequency
  int32_t D, B;      // dishes, beams
  int32_t bg;        // beam group (NT 32-beam tiles)
  int32_t jlayout;   // 0 = [b][f][p][t], 1 = [b][f][t][p]
};

// Stores the 16 bytes (v.x .. v.w, little-endian) at dst, keeping bytes
// i < n. whole: every piece of the launch has n >= 16 (uniform).
__device__ __forceinline__ void store_piece(uint8_t *dst, const v4i v, int n, bool whole) {
  if (whole)
    *gstore_ptr<v4i_u>(dst) = v;
  else
    store_row16(dst, (uint32_t)v.x, (uint32_t)v.y, (uint32_t)v.z, (uint32_t)v.w, n);
}

// Stores the 4 bytes of w (little-endian) at dst, keeping bytes i < n.
__device__ __forceinline__ void store_row4(uint8_t *dst, uint32_t w, int n, bool whole) {
  if (whole || n >= 4) {
    *gstore_ptr<u32_u>(dst) = w;
  } else {
    BF_GLOBAL uint8_t *d = gstore_ptr<uint8_t>(dst);
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (i < n) d[i] = (uint8_t)(w >> (8 * i));
  }
}

// The tile pipeline of a workgroup: tiles tile_first .. tile_end - 1 of one
// job, in rounds of kWaves consecutive tiles (tile tile_first + kWaves r +
// wave in round r), against the weights staged in LDS (Wl, Cl, Gl). Sl is
// the J staging area (NULL: direct stores). All waves of the workgroup call
// this with the same arguments.
//
// MFMA forms. int8, v_mfma_i32_16x16x64_i8: per K step of 32 dishes, for
// each polarization q (a 16-row block of the tile: its 16 times) and each
// 16-beam column group, two MFMAs (real and imaginary accumulator). Lane
// (kg, col) = (lane / 16, lane % 16) holds, for the E operand, row (time)
// col and dishes 8 kg .. 8 kg + 7 of the step; for the weight operand, beam
// col of the group and the same dishes; its accumulators hold rows
// 4 kg .. 4 kg + 3 (times) of column col (beam). fp6 (KS = 4 .. 32),
// v_mfma_scale_f32_16x16x128_f8f6f4 with E2M3 x E3M2 operands: the dishes
// of two consecutive steps in one MFMA (see pack_e_f6 and pack_weights_f6),
// fp32 accumulators of exact integer sums, and no int-to-float conversion in
// the epilogue. The launch runs at the board's power limit, so energy per
// operation sets its time: the fp6 form took the full-size launch from
// 314 to 286 us against the same instruction with fp8 operands (E4M3 x
// E5M2), which had taken it from 343 us with v_mfma_i32_16x16x64_i8.
template <int NT, int KS, bool DT, int JL>
__device__ __forceinline__ void run_tiles(const TileJob &job, const int64_t tile_first, const int64_t tile_end,
                                          const v4i *Wl, const int2 *Cl, const float *Gl, v4i *Sl) {
  static_assert(NT >= 1 && NT <= kMaxBeamTiles, "32-beam tiles per workgroup");
  static_assert(KS >= 2 && KS <= 64 && KS % 2 == 0, "K steps of 16 dishes, in pairs");
  constexpr int KQ = KS / 2;                 // K steps of 32 dishes (one E load each)
  constexpr bool F6 = mfma_f6(KS);           // fp6 16x16x128 MFMA over pairs of steps (64 dishes), fp32 accumulators
  typedef typename std::conditional<F6, v4f, v4i>::type Acc;
  constexpr int NQ = 2 * NT;                 // 16-beam column groups
  constexpr int RP0 = NT >= 3 ? 4 : 8;       // E prefetch distance in steps (fewer registers at NT >= 3)
  constexpr int RP = RP0 < KQ ? RP0 : KQ;
  static_assert(KQ % RP == 0, "ring slots must line up across tiles");
  constexpr int64_t step = kWaves;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t rounds = tile_end > tile_first ? (tile_end - tile_first + step - 1) / step : 0;
  int64_t tile = tile_first + wave;
  // Every J piece is complete (uniform).
  const bool fast_store = (job.T & 15) == 0;

  const int lane = threadIdx.x & 63;
  const int col = lane & 15; // E row (time) of the operand; beam column of the weights and accumulators
  const int kg = lane >> 4;  // dish group of the operands; time group 4 kg .. 4 kg + 3 of the accumulators

  // E loads straight into the operand layout. Row t of this frequency
  // starts at E + (t F + f) D 2. For K step s the operand of lane (kg, col)
  // is bytes 64 s + 16 kg .. + 15 of row col of the tile (dishes 32 s + 8 kg
  // .. + 7, both polarizations): the four lanes kg = 0..3 of a row read its
  // whole 64-byte segment within one load instruction. (Loading coalesced
  // lane quads and transposing through LDS instead costs 3% more time at
  // full size: the LDS round trip spends more energy than the load saves.)
  // Loads are unconditional and in bounds: rows past T re-read row T - 1
  // (their outputs are computed but never stored) and, with DT, dishes past
  // D read the row's first bytes (their weights are zero, so they add 0).
  const int ld_row = col, ld_piece = kg;
  const int64_t e_tstride = job.e_tstride;
  const uint8_t *Ef = job.E + ld_piece * 16;
  auto row_ptr = [&](int64_t tl) -> const uint8_t * { return Ef + e_row_index(tl, ld_row, job.T) * e_tstride; };
  // With a dish tail the per-lane clamp limit is made opaque in every tile
  // (below), so that the clamped offsets are computed at each load instead of
  // being hoisted out of the tile loop and held in registers.
  int e_lim = e_col_limit(ld_piece, job.D);
  auto load_step = [&](const uint8_t *p, int s) -> v4i {
    const int off = DT ? e_col_offset_lim(s, e_lim) : 64 * s;
    return *gload_ptr<v4i_u>(p + off); // any alignment
  };

  // Nothing issued before this call is waited for inside it: completing it
  // here gives the compiler an empty count of memory operations in flight at
  // the loop below, so its waits on the E ring stay exact when the caller
  // itself loops (the batch kernel runs one call per segment).
  __builtin_amdgcn_s_waitcnt(0x0f70); // vmcnt(0)
  v4i ring[RP];
  {
    const uint8_t *p0 = row_ptr(tile);
#pragma unroll
    for (int q = 0; q < RP; ++q) ring[q] = load_step(p0, q);
  }

  for (int64_t r = 0; r < rounds; ++r, tile += step) {
  if (tile < tile_end) { // wave-uniform
    if constexpr (DT) __asm__ volatile("" : "+v"(e_lim));
    Acc acc[2][NQ][2];
#pragma unroll
    for (int qq = 0; qq < NQ; ++qq) {
      const int2 c = Cl[qq * 16 + col];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if constexpr (F6) {
          const float cr = __int_as_float(c.x), ci = __int_as_float(c.y);
          acc[q][qq][0] = v4f{cr, cr, cr, cr};
          acc[q][qq][1] = v4f{ci, ci, ci, ci};
        } else {
          acc[q][qq][0] = v4i{-c.x, -c.x, -c.x, -c.x};
          acc[q][qq][1] = v4i{-c.y, -c.y, -c.y, -c.y};
        }
      }
    }
    const uint8_t *pcur = row_ptr(tile);
    const uint8_t *pnext = row_ptr(tile + step);

    // One K loop in straight-line code, one scheduling region per step,
    // software pipelined: during step s the wave takes the raw E of step
    // s + 1 from the ring (refilling its slot RP steps ahead), issues the
    // MFMAs of step s group by group, each followed by the read of that
    // group's weight operand of step s + 1, then unpacks the E operands of
    // step s + 1 (VALU work grouped after the MFMAs).
    {
      auto operands = [&](v4i e, v4i (&v1)[2], v4i (&v2)[2]) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const v8i u = unpack_e_pol(e, q ? 0x07050301u : 0x06040200u);
          v1[q] = op_v1(u);
          v2[q] = op_v2(u);
        }
      };
      // takes step s out of the ring (into `staged`) and refills its slot
      v4i staged;
      auto stage_step = [&](int s) {
        const int q = s % RP;
        staged = ring[q];
        const int s2 = s + RP;
        ring[q] = s2 >= KQ ? load_step(pnext, s2 - KQ) : load_step(pcur, s2);
      };
      if constexpr (F6) {
        // Steps of 64 dishes: the E loads of the steps 2 h and 2 h + 1 give,
        // per polarization, the two variants v2, v1 (6 dwords each,
        // pack_e_f6) against the group's 6-dword weight operand (dwords 0..3
        // in entry 2 h of the LDS table, 4..5 in entry 2 h + 1).
        constexpr int KH = KQ / 2;
        auto operands6 = [&](v4i e0, v4i e1, v6i (&o2)[2], v6i (&o1)[2]) {
#pragma unroll
          for (int q = 0; q < 2; ++q) pack_e_f6(e0, e1, q ? 0x01010101u : 0u, o2[q], o1[q]);
        };
        auto ext = [](v6i a) -> v8i { return __builtin_shufflevector(a, a, 0, 1, 2, 3, 4, 5, -1, -1); };
        constexpr int WF = NT == 3 ? NQ / 2 : NQ;
        static_assert(NQ % WF == 0, "weight slot of group g is qq % WF");
        constexpr int NGF = KH * NQ;
        auto wload = [&](int g, v4i &w0, v2i &w1) {
          const int hh = g / NQ, qq = g - hh * NQ;
          w0 = Wl[(2 * hh * NQ + qq) * 64 + lane];
          w1 = *reinterpret_cast<const v2i *>(&Wl[((2 * hh + 1) * NQ + qq) * 64 + lane]);
        };
        v6i o2c[2], o1c[2];
        v4i wc0[WF];
        v2i wc1[WF];
        {
          stage_step(0);
          const v4i r0 = staged;
          stage_step(1);
          const v4i r1 = staged;
#pragma unroll
          for (int g = 0; g < WF; ++g) wload(g, wc0[g], wc1[g]);
          operands6(r0, r1, o2c, o1c);
        }
#pragma unroll
        for (int h = 0; h < KH; ++h) {
          v6i o2n[2], o1n[2];
          v4i rn0, rn1;
          if (h + 1 < KH) {
            stage_step(2 * h + 2);
            rn0 = staged;
            stage_step(2 * h + 3);
            rn1 = staged;
          }
#pragma unroll
          for (int qq = 0; qq < NQ; ++qq) {
            const int g = h * NQ + qq;
            const v4i wx = wc0[qq % WF];
            const v2i wy = wc1[qq % WF];
            const v8i w = v8i{wx.x, wx.y, wx.z, wx.w, wy.x, wy.y, 0, 0};
#pragma unroll
            for (int q = 0; q < 2; ++q) {
              acc[q][qq][0] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(ext(o1c[q]), w, acc[q][qq][0], 2, 3, 0,
                                                                               130, 0, 127);
              acc[q][qq][1] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(ext(o2c[q]), w, acc[q][qq][1], 2, 3, 0,
                                                                               130, 0, 127);
            }
            if (g + WF < NGF) wload(g + WF, wc0[qq % WF], wc1[qq % WF]);
            __builtin_amdgcn_sched_barrier(0);
          }
          if (h + 1 < KH) operands6(rn0, rn1, o2n, o1n);
          __builtin_amdgcn_sched_barrier(0);
          if (h + 1 < KH) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
              o2c[q] = o2n[q];
              o1c[q] = o1n[q];
            }
          }
        }
      } else {
      v4i v1c[2], v2c[2], wc[NQ];
      {
        stage_step(0);
        const v4i raw0 = staged;
#pragma unroll
        for (int qq = 0; qq < NQ; ++qq) wc[qq] = Wl[qq * 64 + lane];
        operands(raw0, v1c, v2c);
      }
#pragma unroll
      for (int s = 0; s < KQ; ++s) {
        v4i v1n[2], v2n[2], rawn;
        if (s + 1 < KQ) {
          stage_step(s + 1);
          rawn = staged;
        }
        // group by group: the four MFMAs of a 16-beam group, then that
        // group's weights of the next step into the same registers
#pragma unroll
        for (int qq = 0; qq < NQ; ++qq) {
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            acc[q][qq][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(v1c[q], wc[qq], acc[q][qq][0], 0, 0, 0);
            acc[q][qq][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(v2c[q], wc[qq], acc[q][qq][1], 0, 0, 0);
          }
          if (s + 1 < KQ) wc[qq] = Wl[((s + 1) * NQ + qq) * 64 + lane];
          __builtin_amdgcn_sched_barrier(0); // keeps the loads behind their group's MFMAs (register pressure)
        }
        if (s + 1 < KQ) operands(rawn, v1n, v2n);
        __builtin_amdgcn_sched_barrier(0);
        if (s + 1 < KQ) {
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            v1c[q] = v1n[q];
            v2c[q] = v2n[q];
          }
        }
      }
      } // int8
    }

    // Epilogue. Accumulator element i of lane (kg, col) is time 4 kg + i of
    // the tile, polarization q, beam col of group qq: one requantised
    // ucomplex4 byte per (re, im) pair, four consecutive times per word.
    const int64_t t0 = tile * kTileT;
    const int nvalid = job.T - t0 < kTileT ? (int)(job.T - t0) : kTileT;
    // The beam column as the epilogue sees it: the empty asm makes it opaque,
    // so that the J addresses below are computed here, per tile, instead of
    // being hoisted out of the tile loop and held in registers through the K
    // loop (12 64-bit addresses at NT = 3).
    int cx = col;
    __asm__ volatile("" : "+v"(cx));
    uint32_t wv[2][NQ];
#pragma unroll
    for (int qq = 0; qq < NQ; ++qq) {
      const float gain = Gl[qq * 16 + col];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        uint32_t b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
          b[i] = jbyte_bits(requant_bits(gain, acc[q][qq][0][i]), requant_bits(gain, acc[q][qq][1][i]));
        wv[q][qq] = pack_low_bytes(b[0], b[1], b[2], b[3]);
      }
    }
    if constexpr (JL == 0) {
      if (Sl) {
        // Row (beam qq * 16 + col, polarization q), bytes 4 kg .. 4 kg + 3 of
        // the wave's 16-byte piece of the round's 128 bytes (see stage_word).
        uint32_t *buf = reinterpret_cast<uint32_t *>(Sl);
#pragma unroll
        for (int qq = 0; qq < NQ; ++qq)
#pragma unroll
          for (int q = 0; q < 2; ++q) buf[stage_word(NT, (int)(r & 1), (qq * 16 + cx) * 2 + q, wave, kg)] = wv[q][qq];
      } else {
        // The E loads of the next tile's first steps were issued thousands
        // of cycles ago; completing them here keeps the stores (counted by
        // the same vmcnt) out of every wait of the next tile.
        __builtin_amdgcn_s_waitcnt(0x0f70); // vmcnt(0)
#pragma unroll
        for (int qq = 0; qq < NQ; ++qq) {
          const int b = job.bg * NT * 32 + qq * 16 + cx;
          if (b < job.B)
#pragma unroll
            for (int q = 0; q < 2; ++q)
              store_row4(job.J + j_row_bfpt(b, job.F, job.f, q, job.T) + t0 + 4 * kg, wv[q][qq], nvalid - 4 * kg,
                         fast_store);
        }
      }
    } else {
      // J[b][f][t][p]: interleave the polarizations of the lane's 4 times:
      // 8 bytes = times 4 kg .. 4 kg + 3 x (p0, p1) of beam qq * 16 + col.
      uint32_t lo[NQ], hi[NQ];
#pragma unroll
      for (int qq = 0; qq < NQ; ++qq) {
        lo[qq] = perm_bytes(wv[1][qq], wv[0][qq], 0x05010400u);
        hi[qq] = perm_bytes(wv[1][qq], wv[0][qq], 0x07030602u);
      }
      if (Sl) {
        // Row = beam, 256 bytes per round (8 tiles x 32 bytes); the wave's
        // 32 bytes are 16-byte pieces 2 wave and 2 wave + 1 (see
        // stage_pair_bftp).
        uint8_t *buf = reinterpret_cast<uint8_t *>(Sl);
#pragma unroll
        for (int qq = 0; qq < NQ; ++qq)
          *reinterpret_cast<v2i *>(buf + stage_pair_bftp(NT, (int)(r & 1), qq * 16 + cx, wave, kg)) =
              v2i{(int)lo[qq], (int)hi[qq]};
      } else {
        __builtin_amdgcn_s_waitcnt(0x0f70); // vmcnt(0), as above
#pragma unroll
        for (int qq = 0; qq < NQ; ++qq) {
          const int b = job.bg * NT * 32 + qq * 16 + cx;
          if (b < job.B)
            store_row8(job.J + j_row_bftp(b, job.F, job.f, job.T) + 2 * (t0 + 4 * kg), lo[qq], hi[qq],
                       2 * (nvalid - 4 * kg));
        }
      }
    }
  } // active tile

  if (Sl) {
    // The round's rows are complete once every wave has passed the barrier
    // (its LDS writes done). Thread q writes piece q & 7 of staging row
    // q >> 3: 8 consecutive threads store one row's 128 contiguous bytes.
    // The other buffer is rewritten only after the next round's barrier,
    // which every wave reaches after finishing these reads.
    __builtin_amdgcn_s_waitcnt(0xc07f); // lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
    const int64_t t_round = (tile_first + r * step) * kTileT;
    const int64_t t_end = tile_end * kTileT < job.T ? tile_end * kTileT : job.T;
    __builtin_amdgcn_s_waitcnt(0x0f70); // vmcnt(0), as above
    if constexpr (JL == 0) {
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        const int q = threadIdx.x + kThreads * k;
        const int srow = q >> 3, seg = q & 7;
        const v4i v = *reinterpret_cast<const v4i *>(reinterpret_cast<const uint8_t *>(Sl) +
                                                     stage_piece_byte(NT, (int)(r & 1), srow, seg));
        const int b = job.bg * NT * 32 + (srow >> 1);
        const int64_t t = t_round + 16 * seg;
        if (b < job.B && t < t_end)
          store_piece(job.J + j_row_bfpt(b, job.F, job.f, srow & 1, job.T) + t, v,
                      t_end - t < 16 ? (int)(t_end - t) : 16, fast_store);
      }
    } else {
      // [b][f][t][p]: thread q writes piece q & 15 (8 times, 16 bytes) of
      // beam row q >> 4; 16 consecutive threads store one row's 256 bytes.
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        const int q = threadIdx.x + kThreads * k;
        const int srow = q >> 4, seg = q & 15;
        const v4i v = *reinterpret_cast<const v4i *>(reinterpret_cast<const uint8_t *>(Sl) +
                                                     stage_piece_bftp(NT, (int)(r & 1), srow, seg));
        const int b = job.bg * NT * 32 + srow;
        const int64_t t = t_round + 8 * seg;
        if (b < job.B && t < t_end)
          store_piece(job.J + j_row_bftp(b, job.F, job.f, job.T) + 2 * t, v,
                      t_end - t < 8 ? (int)(2 * (t_end - t)) : 16, fast_store);
      }
    }
  }
  } // rounds
}

template <int NT, int KS, bool DT, int JL>
__global__ __launch_bounds__(kThreads) void form_beams_mfma(const MfmaParams a) {
  extern __shared__ v4i Wl[];
  int2 *Cl = reinterpret_cast<int2 *>(Wl + KS * NT * 64);
  float *Gl = reinterpret_cast<float *>(Cl + NT * 32);
  v4i *Xl = reinterpret_cast<v4i *>(Gl + NT * 32); // the J staging area, where the launch has one

  const int work = work_of_block(blockIdx.x, a.per_xcd);
  if (work >= a.nwork) return; // uniform per workgroup, before any barrier
  const WorkItem wi = decode_work(work, a.nbg, a.nchunks);
  stage_weights<NT, KS>(a.W, a.corr, a.G, a.NTP, a.wrow0 + wi.f, wi.bg, Wl, Cl, Gl);
  __syncthreads();

  const int64_t ntiles = (a.T + kTileT - 1) / kTileT;
  int64_t tile_begin, tile_end;
  chunk_tiles(wi.chunk, a.tiles_per_chunk, ntiles, &tile_begin, &tile_end);
  TileJob job;
  job.E = a.E + (int64_t)wi.f * a.D * 2;
  job.e_tstride = (int64_t)a.F * a.D * 2;
  job.J = a.J;
  job.T = a.T;
  job.F = a.F;
  job.f = wi.f;
  job.D = a.D;
  job.B = a.B;
  job.bg = wi.bg;
  job.jlayout = a.jlayout;
  v4i *Sl = mfma_stages_j(NT, KS) && a.stage_j ? Xl : nullptr; // compiled out where it cannot fit
  run_tiles<NT, KS, DT, JL>(job, tile_begin, tile_end, Wl, Cl, Gl, Sl);
}

// ---------------------------------------------------------------------------
// Batched form_beams: independent chunks (nf consecutive channels each, own
// T) in one persistent launch. The host cuts the work into segments (chunk,
// frequency of the chunk, beam group, tile range) and groups consecutive
// segments into work items and runs an item's segments in order, restaging
// the weights in LDS only when the (channel, beam group) changes. Workgroup
// w runs item w first; balanced items (at most one per workgroup) end
// there, fixed-size items beyond the grid come from a device-side queue.
// The queue is a counter in device memory that only grows: a launch adds
// exactly its dynamically fetched items (nitems - grid, if positive) plus
// one failed fetch per workgroup, so its fetches are the counter values
// base .. and item grid + (value - base), with base written by the host
// beside the item count, and nothing has to reset it. The descriptors (item
// count, base and mode; per item its first segment record at index item
// and the range of its further records) are read straight from pinned host
// memory, so a launch needs no copy before it: a workgroup reads the header,
// its item's first segment record and the range of the others in one round
// trip over the bus (the others, if any, in one more).
struct BatchParams {
  const SegRec *segs;       // segment records: item w's first at w, then the others of every item
  const int *extra;         // item w's further records: segs[extra[2 w] .. extra[2 w + 1] - 1]
  const int *nitems;        // [number of work items, queue base, 1: queue / 0: one item per workgroup]
  unsigned *head;           // device: the queue counter
  const v4i *W;       // channel table [C][KS / 2][2 NTP][64]
  const int *corr;    // [C][BP][2]
  const float *G;     // [C][BP]
  int32_t NTP, D, B, jlayout;
  int32_t stage_j; // 1: J rows go through the LDS staging area (allocated in the launch's LDS)
};

// Segment records a workgroup holds in LDS at a time.
constexpr int kSegBatch = 64;

template <int NT, int KS, bool DT, int JL>
__global__ __launch_bounds__(kThreads) void form_beams_batch(const BatchParams a) {
  extern __shared__ v4i Wl[];
  int2 *Cl = reinterpret_cast<int2 *>(Wl + KS * NT * 64);
  float *Gl = reinterpret_cast<float *>(Cl + NT * 32);
  v4i *Xl = reinterpret_cast<v4i *>(Gl + NT * 32); // the J staging area, where the launch has one
  __shared__ int s_item;
  __shared__ int s_hdr[3];   // item count, queue base, mode
  __shared__ int s_range[2]; // the item's further segment records
  __shared__ SegRec s_seg[kSegBatch];
  v4i *Sl = mfma_stages_j(NT, KS) && a.stage_j ? Xl : nullptr; // compiled out where it cannot fit
  const int tx = threadIdx.x;
  int loaded_channel = -1, loaded_bg = -1;
  int nitems = 0;
  unsigned base = 0;
  bool queue = false;
  // The workgroup's own item first (the host sizes the blocks for every
  // workgroup): consecutive items (pieces of one chunk, the same weights)
  // on one XCD, where they share its L2, when the grid spreads evenly over
  // the 8 XCDs.
  const int own_item = (gridDim.x & 7) == 0 ? work_of_block(blockIdx.x, gridDim.x >> 3) : blockIdx.x;
  for (int round = 0;; ++round) {
    int it = own_item;
    if (round > 0) {
      if (!queue) break;
      if (tx == 0) {
        const unsigned u = atomicAdd(a.head, 1u) - base; // this launch's u-th fetch
        const unsigned left = nitems > (int)gridDim.x ? (unsigned)(nitems - (int)gridDim.x) : 0u;
        s_item = u < left ? (int)gridDim.x + (int)u : nitems;
      }
      __syncthreads();
      it = __builtin_amdgcn_readfirstlane(s_item);
      if (it >= nitems) break; // uniform
    }
    // one round trip: the item's first segment record, the range of its
    // others and, in the first round, the header
    if (tx < 3) reinterpret_cast<v4i *>(s_seg)[tx] = reinterpret_cast<const v4i *>(a.segs + it)[tx];
    if (tx == 64) {
      s_range[0] = a.extra[2 * it];
      s_range[1] = a.extra[2 * it + 1];
    }
    if (round == 0 && tx == 128) {
      s_hdr[0] = a.nitems[0];
      s_hdr[1] = a.nitems[1];
      s_hdr[2] = a.nitems[2];
    }
    __syncthreads();
    if (round == 0) {
      nitems = __builtin_amdgcn_readfirstlane(s_hdr[0]);
      base = (unsigned)__builtin_amdgcn_readfirstlane(s_hdr[1]);
      queue = __builtin_amdgcn_readfirstlane(s_hdr[2]) != 0; // uniform over the launch
      if (it >= nitems) { // no item of its own: the queue's, if any
        if (queue) continue;
        break;
      }
    }
    const int x1 = __builtin_amdgcn_readfirstlane(s_range[1]);
    int next = __builtin_amdgcn_readfirstlane(s_range[0]);
    int nb = 1; // records in s_seg
    for (int batch = 0;; ++batch) {
      // the further records, up to kSegBatch at a time (after the first
      // record in the first batch), in one more round trip
      const int at = batch == 0 ? 1 : 0;
      const int ne = x1 - next < kSegBatch - at ? x1 - next : kSegBatch - at;
      if (ne > 0) {
        if (tx < 3 * ne) reinterpret_cast<v4i *>(s_seg + at)[tx] = reinterpret_cast<const v4i *>(a.segs + next)[tx];
        next += ne;
        nb = at + ne;
        __syncthreads();
      } else if (batch > 0) {
        break;
      }
      for (int k = 0; k < nb; ++k) {
        // wave-uniform copies of the record (scalar control flow below)
        const SegRec &r = s_seg[k];
        auto uni = [](int32_t x) { return __builtin_amdgcn_readfirstlane(x); };
        auto uni64 = [&](int64_t x) {
          return (int64_t)(((uint64_t)(uint32_t)uni((int32_t)(x >> 32)) << 32) | (uint32_t)uni((int32_t)x));
        };
        const int channel = uni(r.channel);
        const int f_bg = uni(r.f_bg);
        const int f = f_bg & 0xffff, bg = f_bg >> 16;
        if (channel != loaded_channel || bg != loaded_bg) {
          stage_weights<NT, KS>(a.W, a.corr, a.G, a.NTP, channel, bg, Wl, Cl, Gl);
          loaded_channel = channel;
          loaded_bg = bg;
        }
        __syncthreads();
        TileJob job;
        job.E = reinterpret_cast<const uint8_t *>(uni64(reinterpret_cast<int64_t>(r.E)));
        const int nf = uni(r.nf);
        job.e_tstride = (int64_t)nf * a.D * 2;
        job.J = reinterpret_cast<uint8_t *>(uni64(reinterpret_cast<int64_t>(r.J)));
        job.T = uni64(r.T);
        job.F = nf;
        job.f = f;
        job.D = a.D;
        job.B = a.B;
        job.bg = bg;
        job.jlayout = a.jlayout;
        run_tiles<NT, KS, DT, JL>(job, uni(r.tile_begin), uni(r.tile_end), Wl, Cl, Gl, Sl);
        __syncthreads(); // the LDS weights, staging and segment records are reused next
      }
      if (next >= x1) break;
    }
    if (queue) __syncthreads(); // s_item is rewritten by the next pop
  }
}

// Kernel tables: one instance per (NT, KS, DT), registered by the
// bf_mfma_ks*.hip translation units. DT = true handles ndishes that are not
// KS * 16 (the last K steps are partly or wholly past D).
typedef void (*MfmaKernel)(const MfmaParams);
typedef void (*BatchKernel)(const BatchParams);

// The instances of one (NT, KS): dish tail or not, J layout 0 ([b][f][p][t])
// or 1 ([b][f][t][p]). The layout is a template parameter so that each
// kernel holds one epilogue and its register budget serves that one alone.
template <int NT, int KS> MfmaKernel mfma_instance(bool dt, int jl) {
  return dt ? (jl ? form_beams_mfma<NT, KS, true, 1> : form_beams_mfma<NT, KS, true, 0>)
            : (jl ? form_beams_mfma<NT, KS, false, 1> : form_beams_mfma<NT, KS, false, 0>);
}
template <int NT, int KS> BatchKernel batch_instance(bool dt, int jl) {
  return dt ? (jl ? form_beams_batch<NT, KS, true, 1> : form_beams_batch<NT, KS, true, 0>)
            : (jl ? form_beams_batch<NT, KS, false, 1> : form_beams_batch<NT, KS, false, 0>);
}

} // namespace bf

// bf_mfma_kernel.hpp -- the gfx950 form-beams kernel (int8 MFMA, exact).
//
// Computes J for one (frequency, beam group, time chunk) per workgroup:
//  * the prepared weight operands of the frequency's beam group (KS x NT
//    KiB) are staged once into LDS and shared by the 8 waves;
//  * each wave owns 16-sample time tiles; it streams the tile's packed E rows
//    (1 byte per dish and polarization, 16 coalesced bytes per lane per K
//    step) from HBM into a register ring that runs R K-steps ahead, splits
//    them into nibble operands in registers, and runs
//    v_mfma_i32_16x16x64_i8 against every beam tile in LDS: 4 MFMAs per
//    (K step, beam tile) = 2 polarizations x (re, im) variants;
//  * the int32 accumulators start at minus the per-beam nibble offsets, so
//    they end at the exact complex sums of cpu.cxx:24-35; the fused epilogue
//    applies the float gain, round-to-nearest-even and the +-7 clamp
//    (cpu.cxx:37-39), packs ucomplex4 bytes and stores J.
// Rows past T and dishes past D are masked loads (zeros); the weights are
// zero there, so they contribute nothing. The plan routes gains outside the
// fast requantisation range (non-finite, or |G * raw| >= 2^31 possible) to
// the generic kernel, which implements the exact x86 lrint semantics.
#pragma once

#include "bf_device.hpp"

namespace bf {

template <int NT, int KS>
__global__ __launch_bounds__(kThreads) void form_beams_mfma(const MfmaParams a) {
  static_assert(NT >= 1 && NT <= 8, "beam tiles per workgroup");
  static_assert(KS >= 1 && KS <= 32, "K steps");
  constexpr int R = KS < 8 ? KS : 8; // E prefetch distance in K steps
  extern __shared__ v4i Wl[];         // [KS][NT][64]

  const int work = work_of_block(blockIdx.x, a.per_xcd);
  if (work >= a.nwork) return; // uniform per workgroup, before any barrier
  const WorkItem wi = decode_work(work, a.nbg, a.nchunks);
  const int f = wi.f;

  // Stage the weight operands of (f, beam group) into LDS.
  {
    const v4i *src = a.W + ((size_t)f * KS * a.NTP + (size_t)wi.bg * NT) * 64;
    for (int i = threadIdx.x; i < KS * NT * 64; i += kThreads) {
      const int s = i / (NT * 64);
      const int rem = i - s * (NT * 64);
      Wl[i] = src[(size_t)s * a.NTP * 64 + rem];
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4;  // k group of this lane in the operands
  const int r = lane & 15;  // E row (time) in the A operand; beam column in D

  const int64_t ntiles = (a.T + kTileT - 1) / kTileT;
  const int64_t tile_begin = (int64_t)wi.chunk * a.tiles_per_chunk;
  const int64_t tile_end = ntiles < tile_begin + a.tiles_per_chunk ? ntiles : tile_begin + a.tiles_per_chunk;
  int64_t tile = tile_begin + wave;
  if (tile >= tile_end) return; // no barrier follows

  const int BP = a.NTP * 16;
  float gain[NT];
  int cre[NT], cim[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int bp = (wi.bg * NT + nt) * 16 + r;
    gain[nt] = a.G[(size_t)f * BP + bp];
    cre[nt] = a.corr[((size_t)f * BP + bp) * 2 + 0];
    cim[nt] = a.corr[((size_t)f * BP + bp) * 2 + 1];
  }

  // E row t of this frequency starts at E + (t F + f) D 2; this lane's 16
  // bytes of K step s are dishes 32 s + 8 g .. +7, i.e. bytes 64 s + 16 g.
  const int64_t e_tstride = (int64_t)a.F * a.D * 2;
  const uint8_t *Ef = a.E + (int64_t)f * a.D * 2 + g * 16;
  auto load_e = [&](int64_t tl, int s) -> v4i {
    const int64_t t = tl * kTileT + r;
    v4i v = {0, 0, 0, 0};
    if (t < a.T && kDishesPerStep * s + 8 * g < a.D)
      v = *reinterpret_cast<const v4i *>(Ef + t * e_tstride + 64 * s);
    return v;
  };

  v4i ring[R];
#pragma unroll
  for (int q = 0; q < R; ++q) ring[q] = load_e(tile, q);

  const bool aligned4 = (a.T & 3) == 0;
  for (; tile < tile_end; tile += kWaves) {
    v4i acc[NT][4];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      acc[nt][0] = v4i{-cre[nt], -cre[nt], -cre[nt], -cre[nt]};
      acc[nt][1] = v4i{-cim[nt], -cim[nt], -cim[nt], -cim[nt]};
      acc[nt][2] = acc[nt][0];
      acc[nt][3] = acc[nt][1];
    }

    // K loop: rounds of R steps; within a round the ring slot is static.
#pragma unroll 1
    for (int s0 = 0; s0 < KS; s0 += R) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int s = s0 + q;
        const v4i e = ring[q];
        {
          const int s2 = s + R;
          const int64_t tl2 = s2 >= KS ? tile + kWaves : tile;
          const int ss2 = s2 >= KS ? s2 - KS : s2;
          if (tl2 < tile_end) ring[q] = load_e(tl2, ss2);
        }
        v4i op[4];
        unpack_e(e, op);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const v4i w = Wl[(s * NT + nt) * 64 + lane];
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[nt][v] = __builtin_amdgcn_mfma_i32_16x16x64_i8(op[v], w, acc[nt][v], 0, 0, 0);
        }
      }
    }

    // Epilogue. Lane (g, r) holds output rows 4g..4g+3 (times) of column r
    // (beam) for every (beam tile, polarization, variant).
    const int64_t t0 = tile * kTileT + 4 * g;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int b = (wi.bg * NT + nt) * 16 + r;
      if (b >= a.B) continue;
      float byte[2][4];
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float qr = requant_fast(gain[nt], acc[nt][2 * p][j]);
          const float qi = requant_fast(gain[nt], acc[nt][2 * p + 1][j]);
          byte[p][j] = jbyte_f(qr, qi);
        }
      const bool full = aligned4 && t0 + 3 < a.T;
      if (a.jlayout == 0) {
        // J[b][f][p][t]
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          uint8_t *dst = a.J + (((int64_t)b * a.F + f) * 2 + p) * a.T + t0;
          if (full) {
            uint32_t w = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) w = put_byte(byte[p][j], j, w);
            *reinterpret_cast<uint32_t *>(dst) = w;
          } else {
#pragma unroll
            for (int j = 0; j < 4; ++j)
              if (t0 + j < a.T) dst[j] = (uint8_t)(int)byte[p][j];
          }
        }
      } else {
        // J[b][f][t][p]
        uint8_t *dst = a.J + (((int64_t)b * a.F + f) * a.T + t0) * 2;
        if (full) {
          uint32_t w0 = 0, w1 = 0;
          w0 = put_byte(byte[0][0], 0, w0);
          w0 = put_byte(byte[1][0], 1, w0);
          w0 = put_byte(byte[0][1], 2, w0);
          w0 = put_byte(byte[1][1], 3, w0);
          w1 = put_byte(byte[0][2], 0, w1);
          w1 = put_byte(byte[1][2], 1, w1);
          w1 = put_byte(byte[0][3], 2, w1);
          w1 = put_byte(byte[1][3], 3, w1);
          *reinterpret_cast<uint2 *>(dst) = make_uint2(w0, w1);
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (t0 + j < a.T) {
              dst[2 * j + 0] = (uint8_t)(int)byte[0][j];
              dst[2 * j + 1] = (uint8_t)(int)byte[1][j];
            }
        }
      }
    }
  }
}

// Kernel table: one instance per (NT, KS) pair, registered by the
// bf_mfma_ks*.hip translation units.
typedef void (*MfmaKernel)(const MfmaParams);
MfmaKernel mfma_kernel_for(int nt, int ks);

} // namespace bf

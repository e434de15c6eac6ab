// bf_kernels_misc.hip -- weight preparation, the shape-generic form-beams
// kernel, device Adler-32 and the seeded input generator.
#include "bf_kernels.hpp"

namespace bf {

// ---------------------------------------------------------------------------
// Weight operands W[f][s][g][lane] (16 bytes each). g is a 16-beam group (2
// NTP of them over the NTP 32-beam tiles) and lane = (kg, c) = (lane / 16,
// lane % 16) serves beam 16 g + c; zero outside B and D. KS counts 16-dish
// steps (even); s runs over the KS / 2 steps of 32 dishes.
//  * int8 (16x16x64 MFMA): entry s holds dishes 32 s + 8 kg .. +7, Ar then
//    Ai (pack_weights).
//  * fp6 (16x16x128 MFMA, KS = 4 .. 32): entries 2 h and 2 h + 1 hold the
//    6-dword operand of the 64-dish step h, over dishes 64 h + 8 kg .. +7
//    and 64 h + 32 + 8 kg .. +7: dwords 0..3 in entry 2 h, 4..5 in entry
//    2 h + 1 (pack_weights_f6).
__global__ void prep_weights_kernel(const uint8_t *__restrict__ A, v4i *__restrict__ W, int F, int B, int D, int KS,
                                    int NTP) {
  const int KQ = KS / 2, NQP = 2 * NTP;
  const int64_t total = (int64_t)F * KQ * NQP * 64;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(i & 63);
    int64_t q = i >> 6;
    const int g = (int)(q % NQP);
    q /= NQP;
    const int s = (int)(q % KQ);
    const int f = (int)(q / KQ);
    const int b = g * 16 + (lane & 15);
    auto gather = [&](int d0, uint8_t a[8]) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int d = d0 + k;
        a[k] = (b < B && d < D) ? A[((int64_t)f * B + b) * D + d] : (uint8_t)0x88; // 0x88 decodes to (0, 0)
      }
    };
    uint8_t a0[8], a1[8];
    if (mfma_f6(KS)) {
      const int h = s >> 1;
      gather(2 * kDishesPerStep * (2 * h) + 8 * (lane >> 4), a0);
      gather(2 * kDishesPerStep * (2 * h + 1) + 8 * (lane >> 4), a1);
      uint32_t w[6];
      pack_weights_f6(a0, a1, w);
      W[i] = (s & 1) ? v4i{(int)w[4], (int)w[5], 0, 0} : v4i{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
    } else {
      gather(2 * kDishesPerStep * s + 8 * (lane >> 4), a0);
      W[i] = pack_weights(a0);
    }
  }
}

// Per-(f, padded beam) nibble-offset corrections and padded gains.
__global__ void prep_corr_kernel(const uint8_t *__restrict__ A, const float *__restrict__ G, int *__restrict__ corr,
                                 float *__restrict__ Gp, int F, int B, int D, int BP) {
  const int64_t total = (int64_t)F * BP;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int f = (int)(i / BP);
    const int b = (int)(i % BP);
    int sr = 0, si = 0;
    float g = 0.0f;
    if (b < B) {
      const uint8_t *row = A + ((int64_t)f * B + b) * D;
      for (int d = 0; d < D; ++d) {
        sr += dec_re(row[d]);
        si += dec_im(row[d]);
      }
      g = G[(int64_t)f * B + b];
    }
    int cr, ci;
    corrections(sr, si, &cr, &ci);
    corr[2 * i + 0] = cr;
    corr[2 * i + 1] = ci;
    Gp[i] = g;
  }
}

// ---------------------------------------------------------------------------
// Shape-generic form_beams: one thread per output, direct dish loop as in
// cpu.cxx:24-35, exact requantisation. Used for shapes outside the MFMA
// kernel's range (P != 2, D not a multiple of 8, D > 1024) and for gains
// outside the fast requantisation range.
__global__ void form_beams_generic_kernel(const uint8_t *__restrict__ E, uint8_t *__restrict__ J,
                                          const uint8_t *__restrict__ A, const float *__restrict__ G, int64_t T,
                                          int F, int D, int P, int B, int jlayout) {
  const int64_t total = (int64_t)B * F * P * T;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    // i enumerates (b, f, p, t) with t fastest.
    const int64_t t = i % T;
    int64_t q = i / T;
    const int p = (int)(q % P);
    q /= P;
    const int f = (int)(q % F);
    const int b = (int)(q / F);
    const uint8_t *Arow = A + ((int64_t)f * B + b) * D;
    const uint8_t *Erow = E + ((t * F + f) * (int64_t)D) * P + p;
    int rr = 0, ri = 0;
    for (int d = 0; d < D; ++d) {
      const uint32_t av = Arow[d], ev = Erow[(int64_t)d * P];
      const int ar = dec_re(av), ai = dec_im(av), er = dec_re(ev), ei = dec_im(ev);
      rr += ar * er - ai * ei;
      ri += ar * ei + ai * er;
    }
    const float g = G[(int64_t)f * B + b];
    const float qr = requant_exact(g, rr), qi = requant_exact(g, ri);
    const int64_t off = jlayout == 0 ? (((int64_t)b * F + f) * P + p) * T + t : (((int64_t)b * F + f) * T + t) * P + p;
    J[off] = (uint8_t)enc((int)qr, (int)qi);
  }
}

// ---------------------------------------------------------------------------
// Adler-32 (adler32.c:7-20) in closed form: with bytes x_i, i = 0..n-1,
//   a = 1 + sum x_i,   b = n + sum (n - i) x_i   (mod 65521).
// Each thread folds 16-byte vectors into (sum x, sum (n - i) x) mod 65521;
// each block writes its pair, reduced mod 65521, to partials[2 block ..];
// adler32_final_kernel folds the pairs (in block order) into the checksum.
// No atomics and no zeroing: every pair is written by its block.

__global__ void adler32_partial_kernel(const uint8_t *__restrict__ data, uint64_t n,
                                       unsigned long long *__restrict__ partials) {
  // Bytes before the first 16-byte boundary and after the last full vector
  // are folded in by thread 0 of block 0.
  const uint64_t pre = ((16u - ((uintptr_t)data & 15u)) & 15u) < n ? ((16u - ((uintptr_t)data & 15u)) & 15u) : n;
  const uint64_t nvec = (n - pre) / 16;
  const uint4 *vec = reinterpret_cast<const uint4 *>(data + pre);
  uint64_t sa = 0, sb = 0;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvec; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint4 x = vec[v];
    const uint32_t w[4] = {x.x, x.y, x.z, x.w};
    uint32_t s1 = 0, s2 = 0; // sum x_j and sum j x_j over the 16 bytes
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t byte = (w[k] >> (8 * j)) & 0xffu;
        s1 += byte;
        s2 += (uint32_t)(4 * k + j) * byte;
      }
    const uint64_t i0 = pre + v * 16;
    const uint64_t wgt = (n - i0) % kAdlerMod; // sum (n - i0 - j) x_j = (n - i0) s1 - s2
    sa += s1;
    sb += (wgt * s1 + (uint64_t)kAdlerMod * 32768u - s2) % kAdlerMod; // s2 <= 30600 < 65521 * 32768
    if (sa >= (1ull << 40)) sa %= kAdlerMod;
    if (sb >= (1ull << 40)) sb %= kAdlerMod;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (uint64_t i = 0; i < pre; ++i) {
      sa += data[i];
      sb += ((n - i) % kAdlerMod) * data[i];
    }
    for (uint64_t i = pre + nvec * 16; i < n; ++i) {
      sa += data[i];
      sb += ((n - i) % kAdlerMod) * data[i];
    }
  }
  sa %= kAdlerMod;
  sb %= kAdlerMod;
  for (int off = 32; off > 0; off >>= 1) { // each term < 65521: no overflow
    sa += __shfl_down(sa, off, 64);
    sb += __shfl_down(sb, off, 64);
  }
  __shared__ uint64_t wsum[2][kAdlerThreads / 64];
  if ((threadIdx.x & 63) == 0) {
    wsum[0][threadIdx.x >> 6] = sa;
    wsum[1][threadIdx.x >> 6] = sb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t a = 0, b = 0;
    for (int w = 0; w < kAdlerThreads / 64; ++w) {
      a += wsum[0][w];
      b += wsum[1][w];
    }
    partials[2 * blockIdx.x + 0] = a % kAdlerMod;
    partials[2 * blockIdx.x + 1] = b % kAdlerMod;
  }
}

// One block: folds nblocks pairs (each < 65521) into the checksum of n bytes.
__global__ void adler32_final_kernel(const unsigned long long *__restrict__ partials, int nblocks, uint64_t n,
                                     uint32_t *__restrict__ out) {
  uint64_t a = 0, b = 0;
  for (int i = threadIdx.x; i < nblocks; i += blockDim.x) {
    a += partials[2 * i + 0];
    b += partials[2 * i + 1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  __shared__ uint64_t wsum[2][kAdlerThreads / 64];
  if ((threadIdx.x & 63) == 0) {
    wsum[0][threadIdx.x >> 6] = a;
    wsum[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t ta = 0, tb = 0;
    for (int w = 0; w < kAdlerThreads / 64; ++w) {
      ta += wsum[0][w];
      tb += wsum[1][w];
    }
    const uint64_t ra = (1 + ta) % kAdlerMod;
    const uint64_t rb = (n % kAdlerMod + tb) % kAdlerMod;
    out[0] = (uint32_t)((rb << 16) | ra);
  }
}

// Blocks of the partial pass for n bytes at data (at most kAdlerBlocks).
int adler32_blocks(const uint8_t *data, uint64_t n) {
  const uint64_t pre = (16u - ((uintptr_t)data & 15u)) & 15u;
  const uint64_t nvec = n > pre ? (n - pre) / 16 : 0;
  const uint64_t b = (nvec + kAdlerThreads - 1) / kAdlerThreads;
  return (int)(b < 1 ? 1 : (b > (uint64_t)kAdlerBlocks ? (uint64_t)kAdlerBlocks : b));
}

// ---------------------------------------------------------------------------
// Seeded uniform ucomplex4 bytes (same definition as the host generator).
BF_HD uint64_t splitmix64_at(uint64_t seed, uint64_t counter) {
  uint64_t z = seed + (counter + 1u) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

BF_HD uint32_t uniform_byte(uint64_t seed, uint64_t counter) {
  const uint64_t z = splitmix64_at(seed, counter);
  const int re = (int)(((z & 0xffffffffull) * 15u) >> 32) - 7;
  const int im = (int)(((z >> 32) * 15u) >> 32) - 7;
  return enc(re, im);
}

__global__ void fill_uniform_kernel(uint8_t *__restrict__ buf, uint64_t n, uint64_t seed, uint64_t offset) {
  // Bytes before the first 16-byte boundary and after the last full vector
  // are written one by one by the first 32 threads of block 0.
  const uint64_t mis = (16u - ((uintptr_t)buf & 15u)) & 15u;
  const uint64_t pre = mis < n ? mis : n;
  const uint64_t nvec = (n - pre) / 16;
  uint4 *vec = reinterpret_cast<uint4 *>(buf + pre);
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvec; v += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      uint32_t x = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) x |= uniform_byte(seed, offset + pre + v * 16 + 4 * k + j) << (8 * j);
      w[k] = x;
    }
    vec[v] = make_uint4(w[0], w[1], w[2], w[3]);
  }
  if (blockIdx.x == 0 && threadIdx.x < 32) {
    const uint64_t k = threadIdx.x;
    const uint64_t i = k < 16 ? k : pre + nvec * 16 + (k - 16);
    const bool in_range = k < 16 ? (i < pre) : (i < n);
    if (in_range) buf[i] = (uint8_t)uniform_byte(seed, offset + i);
  }
}

} // namespace bf

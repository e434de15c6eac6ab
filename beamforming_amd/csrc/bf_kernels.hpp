// bf_kernels.hpp -- declarations of every kernel in the library.
#pragma once

#include "bf_mfma_kernel.hpp"

namespace bf {

__global__ void prep_weights_kernel(const uint8_t *__restrict__ A, v4i *__restrict__ W, int F, int B, int D, int KS,
                                    int NTP);
__global__ void prep_corr_kernel(const uint8_t *__restrict__ A, const float *__restrict__ G, int *__restrict__ corr,
                                 float *__restrict__ Gp, int F, int B, int D, int BP);
__global__ void form_beams_generic_kernel(const uint8_t *__restrict__ E, uint8_t *__restrict__ J,
                                          const uint8_t *__restrict__ A, const float *__restrict__ G, int64_t T,
                                          int F, int D, int P, int B, int jlayout);
// Device Adler-32: kAdlerThreads threads per block, at most kAdlerBlocks
// blocks; the scratch of one checksum is kAdlerScratchWords 64-bit words
// (2 per block, then the 32-bit result).
constexpr uint32_t kAdlerMod = 65521u;
constexpr int kAdlerThreads = 256;
constexpr int kAdlerBlocks = 2048;
constexpr int kAdlerScratchWords = 2 * kAdlerBlocks + 2;
__global__ void adler32_partial_kernel(const uint8_t *__restrict__ data, uint64_t n,
                                       unsigned long long *__restrict__ partials);
__global__ void adler32_final_kernel(const unsigned long long *__restrict__ partials, int nblocks, uint64_t n,
                                     uint32_t *__restrict__ out);
int adler32_blocks(const uint8_t *data, uint64_t n);
__global__ void fill_uniform_kernel(uint8_t *__restrict__ buf, uint64_t n, uint64_t seed, uint64_t offset);

// MFMA kernel instances, one translation unit per K-step count.
MfmaKernel mfma_kernel_ks2(int nt, bool dish_tail, int jl);
MfmaKernel mfma_kernel_ks4(int nt, bool dish_tail, int jl);
MfmaKernel mfma_kernel_ks8(int nt, bool dish_tail, int jl);
MfmaKernel mfma_kernel_ks16(int nt, bool dish_tail, int jl);
MfmaKernel mfma_kernel_ks32(int nt, bool dish_tail, int jl);
MfmaKernel mfma_kernel_ks64(int nt, bool dish_tail, int jl);
BatchKernel batch_kernel_ks2(int nt, bool dish_tail, int jl);
BatchKernel batch_kernel_ks4(int nt, bool dish_tail, int jl);
BatchKernel batch_kernel_ks8(int nt, bool dish_tail, int jl);
BatchKernel batch_kernel_ks16(int nt, bool dish_tail, int jl);
BatchKernel batch_kernel_ks32(int nt, bool dish_tail, int jl);
BatchKernel batch_kernel_ks64(int nt, bool dish_tail, int jl);

} // namespace bf

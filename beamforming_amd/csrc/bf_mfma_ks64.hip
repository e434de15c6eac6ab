// bf_mfma_ks64.hip -- MFMA form-beams kernel instances with 64 K steps (1024 dishes):
// the plain kernel and the persistent batch kernel.
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks64(int nt, bool dish_tail, int jl) {
  switch (nt) {
  case 1:
    return mfma_instance<1, 64>(dish_tail, jl);
  case 2:
    return mfma_instance<2, 64>(dish_tail, jl);
  default:
    return nullptr;
  }
}

BatchKernel batch_kernel_ks64(int nt, bool dish_tail, int jl) {
  switch (nt) {
  case 1:
    return batch_instance<1, 64>(dish_tail, jl);
  case 2:
    return batch_instance<2, 64>(dish_tail, jl);
  default:
    return nullptr;
  }
}

} // namespace bf

// bf_mfma_ks8.hip -- MFMA form-beams kernel instances with 8 K steps (128 dishes):
// the plain kernel and the persistent batch kernel.
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks8(int nt, bool dish_tail, int jl) {
  switch (nt) {
  case 1:
    return mfma_instance<1, 8>(dish_tail, jl);
  case 2:
    return mfma_instance<2, 8>(dish_tail, jl);
  case 3:
    return mfma_instance<3, 8>(dish_tail, jl);
  default:
    return nullptr;
  }
}

BatchKernel batch_kernel_ks8(int nt, bool dish_tail, int jl) {
  switch (nt) {
  case 1:
    return batch_instance<1, 8>(dish_tail, jl);
  case 2:
    return batch_instance<2, 8>(dish_tail, jl);
  case 3:
    return batch_instance<3, 8>(dish_tail, jl);
  default:
    return nullptr;
  }
}

} // namespace bf

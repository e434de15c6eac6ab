// bf_mfma_ks16.hip -- MFMA form-beams kernel instances with 16 K steps (256 dishes):
// the plain kernel and the persistent batch kernel.
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks16(int nt, bool dish_tail, int jl) {
  switch (nt) {
  case 1:
    return mfma_instance<1, 16>(dish_tail, jl);
  case 2:
    return mfma_instance<2, 16>(dish_tail, jl);
  case 3:
    return mfma_instance<3, 16>(dish_tail, jl);
  default:
    return nullptr;
  }
}

BatchKernel batch_kernel_ks16(int nt, bool dish_tail, int jl) {
  switch (nt) {
  case 1:
    return batch_instance<1, 16>(dish_tail, jl);
  case 2:
    return batch_instance<2, 16>(dish_tail, jl);
  case 3:
    return batch_instance<3, 16>(dish_tail, jl);
  default:
    return nullptr;
  }
}

} // namespace bf

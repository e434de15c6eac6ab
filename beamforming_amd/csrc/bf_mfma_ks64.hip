// bf_mfma_ks64.hip -- MFMA form-beams kernel instances with 64 K steps (1024 dishes):
// the plain kernel and the persistent batch kernel.
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks64(int nt, bool dish_tail) {
  switch (nt) {
  case 1:
    return dish_tail ? form_beams_mfma<1, 64, true> : form_beams_mfma<1, 64, false>;
  case 2:
    return dish_tail ? form_beams_mfma<2, 64, true> : form_beams_mfma<2, 64, false>;
  default:
    return nullptr;
  }
}

BatchKernel batch_kernel_ks64(int nt, bool dish_tail) {
  switch (nt) {
  case 1:
    return dish_tail ? form_beams_batch<1, 64, true> : form_beams_batch<1, 64, false>;
  case 2:
    return dish_tail ? form_beams_batch<2, 64, true> : form_beams_batch<2, 64, false>;
  default:
    return nullptr;
  }
}

} // namespace bf

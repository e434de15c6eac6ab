// bf_mfma_ks32.hip -- MFMA form-beams kernel instances with 32 K steps (512 dishes):
// the plain kernel and the persistent batch kernel.
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks32(int nt, bool dish_tail) {
  switch (nt) {
  case 1:
    return dish_tail ? form_beams_mfma<1, 32, true> : form_beams_mfma<1, 32, false>;
  case 2:
    return dish_tail ? form_beams_mfma<2, 32, true> : form_beams_mfma<2, 32, false>;
  case 3:
    return dish_tail ? form_beams_mfma<3, 32, true> : form_beams_mfma<3, 32, false>;
  case 4:
    return dish_tail ? form_beams_mfma<4, 32, true> : form_beams_mfma<4, 32, false>;
  default:
    return nullptr;
  }
}

BatchKernel batch_kernel_ks32(int nt, bool dish_tail) {
  switch (nt) {
  case 1:
    return dish_tail ? form_beams_batch<1, 32, true> : form_beams_batch<1, 32, false>;
  case 2:
    return dish_tail ? form_beams_batch<2, 32, true> : form_beams_batch<2, 32, false>;
  case 3:
    return dish_tail ? form_beams_batch<3, 32, true> : form_beams_batch<3, 32, false>;
  case 4:
    return dish_tail ? form_beams_batch<4, 32, true> : form_beams_batch<4, 32, false>;
  default:
    return nullptr;
  }
}

} // namespace bf

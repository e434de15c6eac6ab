// bf_mfma_ks16.hip -- MFMA form-beams kernel instances with 16 K steps (256 dishes):
// the plain kernel and the persistent batch kernel.
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks16(int nt, bool dish_tail) {
  switch (nt) {
  case 1:
    return dish_tail ? form_beams_mfma<1, 16, true> : form_beams_mfma<1, 16, false>;
  case 2:
    return dish_tail ? form_beams_mfma<2, 16, true> : form_beams_mfma<2, 16, false>;
  case 3:
    return dish_tail ? form_beams_mfma<3, 16, true> : form_beams_mfma<3, 16, false>;
  case 4:
    return dish_tail ? form_beams_mfma<4, 16, true> : form_beams_mfma<4, 16, false>;
  default:
    return nullptr;
  }
}

BatchKernel batch_kernel_ks16(int nt, bool dish_tail) {
  switch (nt) {
  case 1:
    return dish_tail ? form_beams_batch<1, 16, true> : form_beams_batch<1, 16, false>;
  case 2:
    return dish_tail ? form_beams_batch<2, 16, true> : form_beams_batch<2, 16, false>;
  case 3:
    return dish_tail ? form_beams_batch<3, 16, true> : form_beams_batch<3, 16, false>;
  case 4:
    return dish_tail ? form_beams_batch<4, 16, true> : form_beams_batch<4, 16, false>;
  default:
    return nullptr;
  }
}

} // namespace bf

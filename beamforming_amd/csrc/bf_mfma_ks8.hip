// bf_mfma_ks8.hip -- MFMA form-beams kernel instances with 8 K steps (128 dishes):
// the plain kernel and the persistent batch kernel.
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks8(int nt, bool dish_tail) {
  switch (nt) {
  case 1:
    return dish_tail ? form_beams_mfma<1, 8, true> : form_beams_mfma<1, 8, false>;
  case 2:
    return dish_tail ? form_beams_mfma<2, 8, true> : form_beams_mfma<2, 8, false>;
  case 3:
    return dish_tail ? form_beams_mfma<3, 8, true> : form_beams_mfma<3, 8, false>;
  case 4:
    return dish_tail ? form_beams_mfma<4, 8, true> : form_beams_mfma<4, 8, false>;
  default:
    return nullptr;
  }
}

BatchKernel batch_kernel_ks8(int nt, bool dish_tail) {
  switch (nt) {
  case 1:
    return dish_tail ? form_beams_batch<1, 8, true> : form_beams_batch<1, 8, false>;
  case 2:
    return dish_tail ? form_beams_batch<2, 8, true> : form_beams_batch<2, 8, false>;
  case 3:
    return dish_tail ? form_beams_batch<3, 8, true> : form_beams_batch<3, 8, false>;
  case 4:
    return dish_tail ? form_beams_batch<4, 8, true> : form_beams_batch<4, 8, false>;
  default:
    return nullptr;
  }
}

} // namespace bf

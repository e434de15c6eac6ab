// bf_mfma_ks4.hip -- MFMA form-beams kernel instances with 4 K steps (64 dishes):
// the plain kernel and the persistent batch kernel.
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks4(int nt, bool dish_tail) {
  switch (nt) {
  case 1:
    return dish_tail ? form_beams_mfma<1, 4, true> : form_beams_mfma<1, 4, false>;
  case 2:
    return dish_tail ? form_beams_mfma<2, 4, true> : form_beams_mfma<2, 4, false>;
  case 3:
    return dish_tail ? form_beams_mfma<3, 4, true> : form_beams_mfma<3, 4, false>;
  case 4:
    return dish_tail ? form_beams_mfma<4, 4, true> : form_beams_mfma<4, 4, false>;
  default:
    return nullptr;
  }
}

BatchKernel batch_kernel_ks4(int nt, bool dish_tail) {
  switch (nt) {
  case 1:
    return dish_tail ? form_beams_batch<1, 4, true> : form_beams_batch<1, 4, false>;
  case 2:
    return dish_tail ? form_beams_batch<2, 4, true> : form_beams_batch<2, 4, false>;
  case 3:
    return dish_tail ? form_beams_batch<3, 4, true> : form_beams_batch<3, 4, false>;
  case 4:
    return dish_tail ? form_beams_batch<4, 4, true> : form_beams_batch<4, 4, false>;
  default:
    return nullptr;
  }
}

} // namespace bf

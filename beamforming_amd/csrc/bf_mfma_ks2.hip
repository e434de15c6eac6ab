// bf_mfma_ks2.hip -- MFMA form-beams kernel instances with 2 K steps (32 dishes):
// the plain kernel and the persistent batch kernel.
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks2(int nt, bool dish_tail) {
  switch (nt) {
  case 1:
    return dish_tail ? form_beams_mfma<1, 2, true> : form_beams_mfma<1, 2, false>;
  case 2:
    return dish_tail ? form_beams_mfma<2, 2, true> : form_beams_mfma<2, 2, false>;
  case 3:
    return dish_tail ? form_beams_mfma<3, 2, true> : form_beams_mfma<3, 2, false>;
  case 4:
    return dish_tail ? form_beams_mfma<4, 2, true> : form_beams_mfma<4, 2, false>;
  default:
    return nullptr;
  }
}

BatchKernel batch_kernel_ks2(int nt, bool dish_tail) {
  switch (nt) {
  case 1:
    return dish_tail ? form_beams_batch<1, 2, true> : form_beams_batch<1, 2, false>;
  case 2:
    return dish_tail ? form_beams_batch<2, 2, true> : form_beams_batch<2, 2, false>;
  case 3:
    return dish_tail ? form_beams_batch<3, 2, true> : form_beams_batch<3, 2, false>;
  case 4:
    return dish_tail ? form_beams_batch<4, 2, true> : form_beams_batch<4, 2, false>;
  default:
    return nullptr;
  }
}

} // namespace bf

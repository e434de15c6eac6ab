// bf_mfma_ks4.hip -- MFMA form-beams kernel instances with 4 K steps (128 dishes).
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks4(int nt) {
  switch (nt) {
  case 1:
    return form_beams_mfma<1, 4>;
  case 2:
    return form_beams_mfma<2, 4>;
  case 3:
    return form_beams_mfma<3, 4>;
  case 4:
    return form_beams_mfma<4, 4>;
  case 6:
    return form_beams_mfma<6, 4>;
  case 8:
    return form_beams_mfma<8, 4>;
  default:
    return nullptr;
  }
}

} // namespace bf

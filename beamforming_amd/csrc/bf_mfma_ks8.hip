// bf_mfma_ks8.hip -- MFMA form-beams kernel instances with 8 K steps (256 dishes).
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks8(int nt) {
  switch (nt) {
  case 1:
    return form_beams_mfma<1, 8>;
  case 2:
    return form_beams_mfma<2, 8>;
  case 3:
    return form_beams_mfma<3, 8>;
  case 4:
    return form_beams_mfma<4, 8>;
  case 6:
    return form_beams_mfma<6, 8>;
  case 8:
    return form_beams_mfma<8, 8>;
  default:
    return nullptr;
  }
}

} // namespace bf

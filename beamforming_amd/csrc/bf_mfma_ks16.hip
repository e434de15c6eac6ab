// bf_mfma_ks16.hip -- MFMA form-beams kernel instances with 16 K steps (512 dishes).
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks16(int nt) {
  switch (nt) {
  case 1:
    return form_beams_mfma<1, 16>;
  case 2:
    return form_beams_mfma<2, 16>;
  case 3:
    return form_beams_mfma<3, 16>;
  case 4:
    return form_beams_mfma<4, 16>;
  case 6:
    return form_beams_mfma<6, 16>;
  case 8:
    return form_beams_mfma<8, 16>;
  default:
    return nullptr;
  }
}

} // namespace bf

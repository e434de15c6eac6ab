// bf_mfma_ks1.hip -- MFMA form-beams kernel instances with 1 K steps (32 dishes).
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks1(int nt) {
  switch (nt) {
  case 1:
    return form_beams_mfma<1, 1>;
  case 2:
    return form_beams_mfma<2, 1>;
  case 3:
    return form_beams_mfma<3, 1>;
  case 4:
    return form_beams_mfma<4, 1>;
  case 6:
    return form_beams_mfma<6, 1>;
  case 8:
    return form_beams_mfma<8, 1>;
  default:
    return nullptr;
  }
}

} // namespace bf

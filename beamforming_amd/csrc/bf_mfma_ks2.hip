// bf_mfma_ks2.hip -- MFMA form-beams kernel instances with 2 K steps (64 dishes).
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks2(int nt) {
  switch (nt) {
  case 1:
    return form_beams_mfma<1, 2>;
  case 2:
    return form_beams_mfma<2, 2>;
  case 3:
    return form_beams_mfma<3, 2>;
  case 4:
    return form_beams_mfma<4, 2>;
  case 6:
    return form_beams_mfma<6, 2>;
  case 8:
    return form_beams_mfma<8, 2>;
  default:
    return nullptr;
  }
}

} // namespace bf

// bf_mfma_ks32.hip -- MFMA form-beams kernel instances with 32 K steps (1024 dishes).
#include "bf_kernels.hpp"

namespace bf {

MfmaKernel mfma_kernel_ks32(int nt) {
  switch (nt) {
  case 1:
    return form_beams_mfma<1, 32>;
  case 2:
    return form_beams_mfma<2, 32>;
  case 3:
    return form_beams_mfma<3, 32>;
  case 4:
    return form_beams_mfma<4, 32>;
  default:
    return nullptr;
  }
}

} // namespace bf

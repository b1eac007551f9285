// fftm_inst_f16.hip -- float, 16-element register budget instantiations of the multiple-FFT kernel.
#include "fftm_core.hpp"

namespace vhipk {
kernel_fn fftm_kernel_f16(int kind) {
    switch (kind) {
    case VHIP_C2C: return fftm_kernel<float, 16, VHIP_C2C>;
    case VHIP_R2C: return fftm_kernel<float, 16, VHIP_R2C>;
    case VHIP_C2R: return fftm_kernel<float, 16, VHIP_C2R>;
    }
    return nullptr;
}
}  // namespace vhipk

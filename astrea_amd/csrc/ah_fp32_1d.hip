// 1D sweep kernels of the split-precision path (FP64 values, FP32 coefficients).
#include "ah_inst.h"

namespace ah {
bool kernels_1d_fp32(int mask, int recon, int solver, KernelSet* ks) {
    return select_kernels<double, 1, true>(mask, recon, solver, ks);
}
bool kernels_1d_small_fp32(int mask, int recon, int solver, KernelSet* ks) {
    return mask == 0x13 && select_kernels_1d_small<double, true>(recon, solver, ks);
}
}  // namespace ah

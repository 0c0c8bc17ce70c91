// 2D sweep kernels of the split-precision path (FP64 values, FP32 coefficients).
#include "ah_inst.h"

namespace ah {
bool kernels_2d_fp32(int mask, int recon, int solver, KernelSet* ks) {
    return select_kernels<double, 2, true>(mask, recon, solver, ks);
}
}  // namespace ah

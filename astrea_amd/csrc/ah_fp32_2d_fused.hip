// 2D hydrodynamic sweeps in one launch per sweep, split-precision path.
#include "ah_inst.h"

namespace ah {
bool kernels_2d_fused_fp32(int recon, int solver, KernelSet* ks) {
    return select_kernels_2d_fused<double, true>(recon, solver, ks);
}
}  // namespace ah

// 2D hydrodynamic sweeps in one launch per sweep, double arithmetic.
#include "ah_inst.h"

namespace ah {
bool kernels_2d_fused_fp64(int recon, int solver, KernelSet* ks) {
    return select_kernels_2d_fused<double, false>(recon, solver, ks);
}
}  // namespace ah

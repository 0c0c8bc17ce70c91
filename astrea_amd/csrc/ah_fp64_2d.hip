// 2D sweep kernels instantiated for double arithmetic.
#include "ah_inst.h"

namespace ah {
bool kernels_2d_fp64(int mask, int recon, int solver, KernelSet* ks) {
    return select_kernels<double, 2, false>(mask, recon, solver, ks);
}
}  // namespace ah

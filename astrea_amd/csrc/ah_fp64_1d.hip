// 1D sweep kernels instantiated for double arithmetic.
#include "ah_inst.h"

namespace ah {
bool kernels_1d_fp64(int mask, int recon, int solver, KernelSet* ks) {
    return select_kernels<double, 1, false>(mask, recon, solver, ks);
}
}  // namespace ah

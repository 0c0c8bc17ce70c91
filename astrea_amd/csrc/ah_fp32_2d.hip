// 2D sweep kernels instantiated for float arithmetic.
#include "ah_inst.h"

namespace ah {
bool kernels_2d_fp32(int mask, int recon, int solver, KernelSet* ks) {
    return select_kernels<float, 2>(mask, recon, solver, ks);
}
}  // namespace ah

// 2D hydrodynamic PPM with the Colella (c) and Peterson-Hammett (ph) limiters, FP64.
#include "ah_inst.h"

namespace ah {
bool kernels_2d_author_fp64(int author, int solver, KernelSet* ks) {
    if ((solver != LLF && solver != HLLC) || !kernels_2d_fp64(0x17, PPM, solver, ks)) return false;
    if (author == AUTH_C) set2d_author<AUTH_C>(ks);
    else if (author == AUTH_PH) set2d_author<AUTH_PH>(ks);
    else return false;
    return true;
}
}  // namespace ah

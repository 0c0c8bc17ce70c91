// 1D sweep kernels instantiated for double arithmetic.
#include "ah_inst.h"

namespace ah {
bool kernels_1d_fp64(int mask, int recon, int solver, KernelSet* ks) {
    return select_kernels<double, 1, false>(mask, recon, solver, ks);
}
bool kernels_1d_small_fp64(int mask, int recon, int solver, KernelSet* ks) {
    return mask == 0x13 && select_kernels_1d_small<double, false>(recon, solver, ks);
}
bool kernels_1d_author_fp64(int author, int solver, KernelSet* ks) {
    if (author == AUTH_C && solver == LLF) set1d<double, 0x13, PPM, LLF, false, AUTH_C>(ks);
    else if (author == AUTH_C && solver == HLLC) set1d<double, 0x13, PPM, HLLC, false, AUTH_C>(ks);
    else if (author == AUTH_PH && solver == LLF) set1d<double, 0x13, PPM, LLF, false, AUTH_PH>(ks);
    else if (author == AUTH_PH && solver == HLLC) set1d<double, 0x13, PPM, HLLC, false, AUTH_PH>(ks);
    else return false;
    return true;
}
}  // namespace ah

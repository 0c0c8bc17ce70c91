// 2D hydrodynamic PPM with the Colella (c) and Peterson-Hammett (ph) limiters, FP64.
#include "ah_inst.h"

namespace ah {
bool kernels_2d_author_fp64(int author, int solver, KernelSet* ks) {
    if (author == AUTH_C && solver == LLF) set2d_author<AUTH_C, LLF>(ks);
    else if (author == AUTH_C && solver == HLLC) set2d_author<AUTH_C, HLLC>(ks);
    else if (author == AUTH_PH && solver == LLF) set2d_author<AUTH_PH, LLF>(ks);
    else if (author == AUTH_PH && solver == HLLC) set2d_author<AUTH_PH, HLLC>(ks);
    else return false;
    return true;
}
}  // namespace ah

// Launch-table interface between the C-ABI layer (ah_api.hip) and the
// translation units that instantiate the kernel templates.
#pragma once
#include <hip/hip_runtime.h>

#include "ah_mhd.h"

namespace ah {

using SweepLaunch = void (*)(const SweepArgs&, int nblocks, hipStream_t);
using MhdLaunch = void (*)(const MhdArgs&, int nblocks, hipStream_t);
using CopyLaunch = void (*)(const MhdArgs&, double* dst, int nblocks, hipStream_t);

// Tile shapes are compile-time constants of each instantiation; the host
// builds tile tables from them.
struct KernelSet {
    SweepLaunch sweep1d = nullptr;
    SweepLaunch recon[2] = {nullptr, nullptr};
    SweepLaunch flux[2] = {nullptr, nullptr};
    int ts1 = 0;
    int rts[2] = {0, 0}, rtt[2] = {0, 0};
    int fts[2] = {0, 0}, ftt[2] = {0, 0};
    int real_bytes = 8;
    // 2D MHD constrained transport
    MhdLaunch trecon[2] = {nullptr, nullptr};
    MhdLaunch corner = nullptr, magL = nullptr, bface = nullptr, refine = nullptr;
    CopyLaunch copyB = nullptr;
    int tts = 64, ttt = 8;
};

// dims: 1/2; mask: AH_VARS_*; recon: PCM..WENO7; solver: LLF/HLLC/HLLD.
// Returns false if that combination is not instantiated.
bool kernels_1d_fp64(int mask, int recon, int solver, KernelSet* ks);
bool kernels_2d_fp64(int mask, int recon, int solver, KernelSet* ks);
bool kernels_1d_fp32(int mask, int recon, int solver, KernelSet* ks);
// 1D hydrodynamic PLM/PPM with LLF or HLLC on the 115-cell tiles
bool kernels_1d_small_fp64(int mask, int recon, int solver, KernelSet* ks);
bool kernels_1d_small_fp32(int mask, int recon, int solver, KernelSet* ks);
// 1D / 2D hydrodynamic PPM with author c (1) or ph (2), FP64, LLF or HLLC
bool kernels_1d_author_fp64(int author, int solver, KernelSet* ks);
bool kernels_2d_author_fp64(int author, int solver, KernelSet* ks);
bool kernels_2d_fp32(int mask, int recon, int solver, KernelSet* ks);
// 2D hydrodynamic PLM/PPM with LLF or HLLC, one launch per sweep (recon[] null)
bool kernels_2d_fused_fp64(int recon, int solver, KernelSet* ks);
bool kernels_2d_fused_fp32(int recon, int solver, KernelSet* ks);

template <class K>
inline void launch_sweep(K kernel, const SweepArgs& a, int nblocks, hipStream_t s, int nthreads = 256) {
    if (nblocks <= 0) return;
    hipLaunchKernelGGL(kernel, dim3(nblocks), dim3(nthreads), 0, s, a);
}

}  // namespace ah

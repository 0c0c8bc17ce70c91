// Snapshots of the state as the reference saves them every step
// (astrea.py:47): sim_variables.convert_conservative of the conservative grid
// -- the 4th-order cell conversion for PPM/WENO, the pointwise one otherwise
// (generic.py:250-255) -- transposed by ortho_axis (generic.py:264: the
// identity in 1D, (1, 0, 2) in 2D).  The conversion is the one the sweep
// kernels apply to their input (ho_cell_prim, Laplacians along x then y), so
// the snapshot equals the reference's to the last bit.
#pragma once
#include "ah_kernels.h"

namespace ah {

// Snapshots taken inside the step program, without the host: after a step of
// problem `prob`, when `mode` asks for it (1: every step, 2: steps whose dt
// was clamped to a checkpoint, astrea.py:74-78), the snapshot goes to slot
// ctl[0] % slots of a caller-owned ring and t to times[] of that slot.
// ctl[1] holds the step count of the last step looked at, so a launch after
// which the problem did not step (inactive, failed) captures nothing.
struct CaptureArgs {
    const ProbState* ps;
    double* ring;   // slots x cells x 8 doubles
    double* times;  // slots
    long long* ctl; // [0] snapshots taken, [1] steps at the last look; null: capture off
    long long cells;
    double t_end;   // states past t_end are not saved (the loop of astrea.py:45 ends there)
    int prob, mode, slots, pad;
};

__device__ inline bool capture_due(const CaptureArgs& c) {
    const ProbState& st = c.ps[c.prob];
    return st.steps != c.ctl[1] && (c.mode == 1 || st.ckpt != 0) && st.t <= c.t_end;
}

template <int MASK, int DIM, bool HO>
__global__ __launch_bounds__(256) void snapshot_kernel(const double* X, long long P, long long off, int n0, int n1,
                                                       int bc, double gamma, double* out, CaptureArgs cap) {
    using V = VS<MASK>;
    const long long cells = (long long)n0 * n1;
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells) return;
    if (cap.ctl) {
        if (!capture_due(cap)) return;
        out = cap.ring + (cap.ctl[0] % cap.slots) * cells * 8;
    }
    const int i = (int)(cell / n1), j = (int)(cell % n1);
    auto Q = [&](int ii, int jj) {
        V8<double> q = zero8<double, MASK>();
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (V::live(c)) q.c[c] = X[V::slot(c) * P + off + (long long)ii * n1 + jj];
        return q;
    };
    const V8<double> q0 = Q(i, j);
    V8<double> w;
    if constexpr (HO) {
        const int im = bcmap(i - 1, n0, bc), ip = bcmap(i + 1, n0, bc);
        const V8<double> qxm = Q(im, j), qxp = Q(ip, j);
        V8<double> qym = zero8<double, MASK>(), qyp = qym;
        if constexpr (DIM == 2) {
            qym = Q(i, bcmap(j - 1, n1, bc));
            qyp = Q(i, bcmap(j + 1, n1, bc));
        }
        w = ho_cell_prim<MASK, DIM, false>(q0, qxm, qxp, qym, qyp, c2p<MASK, false>(q0, gamma),
                                            c2p<MASK, false>(qxm, gamma), c2p<MASK, false>(qxp, gamma),
                                            c2p<MASK, false>(qym, gamma), c2p<MASK, false>(qyp, gamma), gamma);
    } else {
        w = c2p<MASK, false>(q0, gamma);
    }
    const long long o = DIM == 2 ? ((long long)j * n0 + i) * 8 : (long long)i * 8;
#pragma unroll
    for (int c = 0; c < 8; ++c) out[o + c] = V::live(c) ? w.c[c] : 0.0;
}

// after the capture launch: record t of the slot and advance the counters
__global__ void capture_commit_kernel(CaptureArgs c) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const ProbState& st = c.ps[c.prob];
    if (st.steps == c.ctl[1]) return;
    const bool due = (c.mode == 1 || st.ckpt != 0) && st.t <= c.t_end;
    c.ctl[1] = st.steps;
    if (!due) return;
    c.times[c.ctl[0] % c.slots] = st.t;
    c.ctl[0] += 1;
}

// start of a capture: nothing taken yet, the current step count already seen
__global__ void capture_init_kernel(CaptureArgs c) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    c.ctl[0] = 0;
    c.ctl[1] = c.ps[c.prob].steps;
}

}  // namespace ah

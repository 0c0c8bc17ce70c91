// Non-template control kernels: time step control (astrea.py:45-85), the
// Runge-Kutta stage combinations (evolvers.py:79-206) and the flux difference
// used by the parity entry point.  Included by ah_api.hip only.
#pragma once
#include "ah_kernels.h"

namespace ah {

// ============================================================================
// Time control (astrea.py:45-85), one thread per problem.
// ============================================================================
__global__ void dt_kernel(const ProbDesc* pd, ProbState* ps, int nprob, int dims) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprob) return;
    ProbState& s = ps[p];
    if (!s.active) return;
    const ProbDesc d = pd[p];
    const double e0 = __longlong_as_double((long long)s.eig[0]);
    const double e1 = dims == 2 ? __longlong_as_double((long long)s.eig[1]) : e0;
    s.eig[0] = s.eig[1] = 0ull;
    if (!(e0 <= 1.7976931348623157e308) || !(e1 <= 1.7976931348623157e308) || s.fail_flag) {
        s.failed = 1;
        s.active = 0;
        return;
    }
    const double a0 = d.dx / e0, a1 = d.dx / e1;
    double dt = d.cfl * (dims == 2 ? (a1 < a0 ? a1 : a0) : a0);
    int hit = 0;
    if (s.t + dt >= d.chkpt * s.idx) {
        dt = d.chkpt * s.idx - s.t;
        s.idx += 1;
        hit = 1;
    }
    s.ckpt = hit;
    s.dt = dt;
}

__global__ void advance_kernel(const ProbDesc* pd, ProbState* ps, int nprob) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprob) return;
    ProbState& s = ps[p];
    if (!s.active) return;
    s.aflags = 0ull;
    const ProbDesc d = pd[p];
    if (s.fail_flag) {
        // the reference raises inside this step (LAPACK on a non-finite
        // Jacobian); the state and time stay those of the last good step
        s.failed = 1;
        s.active = 0;
        return;
    }
    s.t = s.t + s.dt;
    s.steps += 1;
    s.parity ^= 1;
    const bool more_time = (s.t <= d.t_end) && !(s.t == d.t_end);
    const bool more_steps = d.max_steps < 0 || s.steps < d.max_steps;
    s.active = more_time && more_steps;
}

// End of a step whose kernels evaluated dt themselves (step_dt): the state
// updates of dt_kernel and advance_kernel in one launch, with the same
// outcome -- a failure seen at stage 0 (non-finite radii) leaves dt and the
// checkpoint index as they were, a later one records this step's dt.
__device__ inline void finish_problem(const ProbDesc& d, ProbState& s, int dims) {
    s.aflags = 0ull;
    const double e0 = __longlong_as_double((long long)s.eig[0]);
    const double e1 = dims == 2 ? __longlong_as_double((long long)s.eig[1]) : e0;
    if (!(e0 <= 1.7976931348623157e308) || !(e1 <= 1.7976931348623157e308)) {
        s.eig[0] = s.eig[1] = 0ull;
        s.failed = 1;
        s.active = 0;
        return;
    }
    int hit;
    const double dt = step_dt(d, s, dims, &hit);
    s.eig[0] = s.eig[1] = 0ull;
    s.idx += hit;
    s.ckpt = hit;
    s.dt = dt;
    if (s.fail_flag) {
        s.failed = 1;
        s.active = 0;
        return;
    }
    s.t = s.t + s.dt;
    s.steps += 1;
    s.parity ^= 1;
    const bool more_time = (s.t <= d.t_end) && !(s.t == d.t_end);
    const bool more_steps = d.max_steps < 0 || s.steps < d.max_steps;
    s.active = more_time && more_steps;
}

__global__ void finish_kernel(const ProbDesc* pd, ProbState* ps, int nprob, int dims) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprob) return;
    ProbState& s = ps[p];
    if (!s.active) return;
    finish_problem(pd[p], s, dims);
}

__global__ void reset_kernel(const ProbDesc* pd, ProbState* ps, int nprob) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprob) return;
    ProbState s;
    s.t = 0.0;
    s.dt = 0.0;
    s.steps = 0;
    s.idx = 1;
    s.failed = 0;
    s.parity = 0;
    s.eig[0] = s.eig[1] = 0ull;
    s.fail_flag = 0;
    s.ckpt = 0;
    s.aflags = 0ull;
    const ProbDesc d = pd[p];
    s.active = (0.0 <= d.t_end) && !(0.0 == d.t_end) && d.max_steps != 0;
    ps[p] = s;
}

// ============================================================================
// Stage combinations of every Runge-Kutta scheme (evolvers.py:79-206),
// written in the reference's operation order.  regs[0] = U (state), regs[1..4]
// = k registers, regs[5..8] = saved L, L = this stage's L.
// ============================================================================
enum RkOp {
    OP_EULER = 0,
    OP_22_0, OP_22_1,
    OP_33_0, OP_33_1, OP_33_2,
    OP_43_0, OP_43_1, OP_43_2, OP_43_3,
    OP_RK4_0, OP_RK4_1, OP_RK4_2, OP_RK4_3,
    OP_53_0, OP_53_1, OP_53_2, OP_53_3, OP_53_4,
    OP_54_0, OP_54_1, OP_54_2, OP_54_3, OP_54_4,
    OP_104_FIRST, OP_104_ACC1, OP_104_K5, OP_104_K2FIRST, OP_104_ACC2, OP_104_FINAL,
    // SSPRK(10,4) with constrained transport refines the increment before
    // adding it (evolvers.py:89-90, 97-99): K4 := dt/6 L, refined, then added
    OP_104_INC, OP_104_ADD_U, OP_104_ADD_K1, OP_104_ADD_K3, OP_104_ADD_K2
};

struct CombineArgs {
    double* regs[9];
    const double* L;
    const ProbDesc* pd;
    ProbState* ps;
    const Tile* tiles;       // cell tiles: (problem, first cell, cell count)
    const int* ntiles_dev;
    long long plane;
    int nv;
    int op;
    int live_dt;   // 1: dt from step_dt (no dt launch in this step program), 0: the stored dt
    int dims;
    // finish: this combination ends the step, and the block that combines a
    // problem's last tile also ends the problem's step (finish_problem); done
    // counts each problem's combined tiles (reset by that block), ct = cells
    // per tile
    int finish;
    int ct;
    unsigned* done;
};

__global__ __launch_bounds__(256) void combine_kernel(CombineArgs a) {
    const long long n = a.plane;
    const int ntiles = wave_uniform(*a.ntiles_dev);
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const Tile tl = ld_const(a.tiles + ti);
        const ProbState& s = a.ps[tl.prob];
        if (!ld_const(&s.active)) continue;
        if (!ld_const(&s.fail_flag)) {
        double dt;
        if (a.live_dt) {
            int hit;
            dt = step_dt(ld_const(&a.pd[tl.prob]), s, a.dims, &hit);
        } else {
            dt = ld_const(&s.dt);
        }
        const long long base = ld_const(&a.pd[tl.prob].off) + tl.s0;
        for (int c = threadIdx.x; c < tl.t0; c += blockDim.x)
        for (int v = 0; v < a.nv; ++v) {
            const long long j = v * n + base + c;
            double* U = a.regs[0];
            double* K1 = a.regs[1];
            double* K2 = a.regs[2];
            double* K3 = a.regs[3];
            double* K4 = a.regs[4];
            const double L = a.L[j];
            switch (a.op) {
                case OP_EULER: U[j] = U[j] + dt * L; break;
                case OP_22_0: K1[j] = U[j] + dt * L; break;
                case OP_22_1: U[j] = .5 * (U[j] + K1[j] + dt * L); break;
                case OP_33_0: K1[j] = U[j] + dt * L; break;
                case OP_33_1: K2[j] = .25 * (3 * U[j] + K1[j] + dt * L); break;
                case OP_33_2: U[j] = (1.0 / 3.0) * (U[j] + 2 * K2[j] + 2 * dt * L); break;
                case OP_43_0: K1[j] = U[j] + .5 * dt * L; break;
                case OP_43_1: K2[j] = K1[j] + .5 * dt * L; break;
                case OP_43_2: K3[j] = (1.0 / 6.0) * (4 * U[j] + 2 * K2[j] + dt * L); break;
                case OP_43_3: U[j] = K3[j] + .5 * dt * L; break;
                case OP_RK4_0: a.regs[5][j] = L; K1[j] = U[j] + .5 * dt * L; break;
                case OP_RK4_1: a.regs[6][j] = L; K2[j] = U[j] + .5 * dt * L; break;
                case OP_RK4_2: a.regs[7][j] = L; K3[j] = U[j] + dt * L; break;
                case OP_RK4_3:
                    U[j] = U[j] + (1.0 / 6.0) * (dt * (a.regs[5][j] + 2 * a.regs[6][j] + 2 * a.regs[7][j] + L));
                    break;
                case OP_53_0: a.regs[5][j] = L; K1[j] = U[j] + .3772689151171 * dt * L; break;
                case OP_53_1: a.regs[6][j] = L; K2[j] = K1[j] + .3772689151171 * dt * L; break;
                case OP_53_2: K3[j] = .56656131914033 * U[j] + .43343868085967 * K2[j] + .16352294089771 * dt * L; break;
                case OP_53_3:
                    K4[j] = .09299483444413 * U[j] + .0000209036962 * K1[j] + .90698426185967 * K3[j] +
                            .00071997378654 * dt * a.regs[5][j] + .34217696850008 * dt * L;
                    break;
                case OP_53_4:
                    U[j] = .0073613226092 * U[j] + .20127980325145 * K1[j] + .00182955389682 * K2[j] +
                           .78952932024253 * K4[j] +
                           (dt * (.0027771981946 * a.regs[5][j] + .00001567934613 * a.regs[6][j] + .29786487010104 * L));
                    break;
                case OP_54_0: K1[j] = U[j] + .39175222657189 * dt * L; break;
                case OP_54_1: K2[j] = .444370493651235 * U[j] + .555629506348765 * K1[j] + .368410593050371 * dt * L; break;
                case OP_54_2: K3[j] = .620101851488403 * U[j] + .379898148511597 * K2[j] + .251891774271694 * dt * L; break;
                case OP_54_3:
                    a.regs[5][j] = L;
                    K4[j] = .178079954393132 * U[j] + .821920045606868 * K3[j] + .544974750228521 * dt * L;
                    break;
                case OP_54_4:
                    U[j] = .517231671970585 * K2[j] + .096059710526147 * K3[j] + .06369246866629 * dt * a.regs[5][j] +
                           .386708617503269 * K4[j] + .226007483236906 * dt * L;
                    break;
                case OP_104_FIRST: K1[j] = U[j] + (1.0 / 6.0) * dt * L; break;
                case OP_104_ACC1: K1[j] = K1[j] + (1.0 / 6.0) * dt * L; break;
                case OP_104_K5: K3[j] = 3.0 / 5.0 * U[j] + 6.0 / 15.0 * K1[j] + 1.0 / 15.0 * dt * L; break;
                case OP_104_K2FIRST: K2[j] = K3[j] + (1.0 / 6.0) * dt * L; break;
                case OP_104_ACC2: K2[j] = K2[j] + (1.0 / 6.0) * dt * L; break;
                case OP_104_INC: K4[j] = (1.0 / 6.0) * dt * L; break;
                case OP_104_ADD_U: K1[j] = U[j] + K4[j]; break;
                case OP_104_ADD_K1: K1[j] = K1[j] + K4[j]; break;
                case OP_104_ADD_K3: K2[j] = K3[j] + K4[j]; break;
                case OP_104_ADD_K2: K2[j] = K2[j] + K4[j]; break;
                case OP_104_FINAL:
                    U[j] = -11.0 / 35.0 * U[j] + 5.0 / 7.0 * K3[j] + 3.0 / 5.0 * K2[j] + 1.0 / 10.0 * dt * L;
                    break;
            }
        }
        }
        if (a.finish) {
            // every block that combines one of the problem's tiles counts it
            // once its threads have read the time state and written their
            // cells; the block that counts the last one ends the problem's
            // step.  No other block reads this problem's time state after
            // counting, so no fence orders the update against them; the next
            // launch sees it.
            __syncthreads();
            if (threadIdx.x == 0) {
                const ProbDesc d = ld_const(&a.pd[tl.prob]);
                const long long cells = (long long)d.n0 * d.n1;
                const unsigned nt = (unsigned)((cells + a.ct - 1) / a.ct);
                if (atomicAdd(&a.done[tl.prob], 1u) == nt - 1) {
                    a.done[tl.prob] = 0u;
                    finish_problem(d, a.ps[tl.prob], a.dims);
                }
            }
        }
    }
}


// L = -sum_a dF_a/dx from sweep-frame interface fluxes (evolvers.py:41-60)
__global__ void fluxdiff_kernel(const double* F0, const double* F1, double* L, int n0, int n1, int dims, int nv,
                                double dx, long long plane, long long fplane0, long long fplane1) {
    const long long ncell = (long long)n0 * n1;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < ncell; i += (long long)gridDim.x * blockDim.x) {
        const int i0 = (int)(i / n1), i1 = (int)(i % n1);
        for (int v = 0; v < nv; ++v) {
            // axis 0: interfaces i0, i0+1 on line i1 (layout [k][t] with t = i1)
            const double d0 = (F0[v * fplane0 + (long long)(i0 + 1) * n1 + i1] - F0[v * fplane0 + (long long)i0 * n1 + i1]) / dx;
            if (dims == 1) {
                L[v * plane + i] = -d0;
            } else {
                const double d1 =
                    (F1[v * fplane1 + (long long)(i1 + 1) * n0 + i0] - F1[v * fplane1 + (long long)i1 * n0 + i0]) / dx;
                L[v * plane + i] = -(d0 + d1);
            }
        }
    }
}

}  // namespace ah

// 2D MHD constrained transport on gfx950 (num_methods/mag_field.py,
// num_methods/evolvers.py:26-32, 51-58, 63-67, 73-76).
//
// Frames: per sweep AX the interface-state arrays are in the sweep-frame
// layout [t][s] (s along AX).  The reference's "ortho frame" of a sweep is the
// transpose of its sweep frame, i.e. exactly this [t][s] order, so the
// transverse reconstruction runs along t of these arrays.  The corner EMF
// e3U is stored in the frame (x, y) = (i0, i1) of the state.
#pragma once
#include "ah_kernels.h"

namespace ah {

struct MhdArgs {
    const void* src[2];   // per sweep: face values feeding the corners (R, [t][s])
    void* wD[2];          // per sweep: transverse extrapolants (R, [t][s])
    void* wU[2];
    double* e3;           // corner EMF, one plane, frame (x, y)
    double* L;            // stage L (B components overwritten)
    double* U;            // state (stage-0 face-field reassignment)
    const double* K;      // register to refine (B planes)
    double* scratch;      // two planes for the refined B
    double* bsave;        // two planes: B_x, B_y of the state before the stage-0 reassignment
    long long plane;
    const ProbDesc* pd;
    const ProbState* ps;
    const Tile* tiles;    // cell tiles (problem, first cell, count)
    const int* ntiles_dev;
    const Tile* ttiles[2];  // transverse-reconstruction tiles per sweep
    const int* nttiles_dev[2];
    int solver_hll;         // corner speeds: HLL branch (mag_field.py:136-149) or eigenvalues (152-159)
};

// PPM 'mc' along t of a sweep's face values (mag_field.py:11-76, method 'ppm')
template <class R, int MASK, int AX, int TS, int TT>
__global__ __launch_bounds__(256) void trecon_kernel(MhdArgs a) {
    using V = VS<MASK>;
    constexpr int NV = V::NV;
    constexpr int H = 4, TW = TT + 2 * H;
    __shared__ R win[NV][TW][TS];
    const int ntiles = wave_uniform(*a.nttiles_dev[AX]);
    const R* src = reinterpret_cast<const R*>(a.src[AX]);
    R* oD = reinterpret_cast<R*>(a.wD[AX]);
    R* oU = reinterpret_cast<R*>(a.wU[AX]);
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const Tile tile = ld_const(a.ttiles[AX] + ti);
        const ProbDesc pd = ld_const(a.pd + tile.prob);
        if (!ld_const(&a.ps[tile.prob].active)) continue;
        const int ns = AX == 0 ? pd.n0 : pd.n1, nt = AX == 0 ? pd.n1 : pd.n0;
        const int bc = pd.bc, s0 = tile.s0, t0 = tile.t0;
        const long long off = pd.off, P = a.plane;
        for (int i = threadIdx.x; i < TW * TS; i += blockDim.x) {
            const int it = i / TS, is = i % TS;
            const int s = s0 + is;
            if (s >= ns) continue;
            const int gt = bcmap(t0 - H + it, nt, bc);
#pragma unroll
            for (int v = 0; v < NV; ++v) win[v][it][is] = src[v * P + off + (long long)gt * ns + s];
        }
        __syncthreads();
        for (int i = threadIdx.x; i < TT * TS; i += blockDim.x) {
            const int it = i / TS, is = i % TS;
            const int s = s0 + is, t = t0 + it;
            if (s >= ns || t >= nt) continue;
            struct WAcc {
                R (*w)[TW][TS];
                int is, bt, n, bc;
                __device__ int e(int p) const { return bc == EDGE ? clampi(p, 0, n - 1) : p; }
                __device__ V8<R> operator()(int m) const {
                    const int j = e(m) - bt;
                    V8<R> r = zero8<R, MASK>();
#pragma unroll
                    for (int c = 0; c < 8; ++c)
                        if (VS<MASK>::live(c)) r.c[c] = w[VS<MASK>::slot(c)][j][is];
                    return r;
                }
                    } W{win, is, t0 - H, nt, bc};
            V8<R> D, Uu;
            recon_cell<R, MASK, PPM, AX>(W, t, MINMOD, FLAT_OFF, D, Uu);
            const long long o = off + (long long)t * ns + s;
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (V::live(c)) {
                    oD[V::slot(c) * P + o] = D.c[c];
                    oU[V::slot(c) * P + o] = Uu.c[c];
                }
        }
        __syncthreads();
    }
}

// corner speeds from the Roe average of (plus = wD[m+1], minus = wU[m])
template <class R, int MASK>
__device__ inline void corner_speeds(const V8<R>& plus, const V8<R>& minus, int ax, R gamma, bool hll, R& ap, R& am) {
    const V8<R> avg = roe_avg<R, MASK>(plus, minus);
    const R rho = avg.c[0], p = avg.c[4];
    const R vx = ax == 0 ? avg.c[1] : avg.c[2];
    const R Bx = ax == 0 ? avg.c[5] : avg.c[6];
    R cf;
    if (hll) {
        // mag_field.py:138-145
        const R cs = ah_sqrt(gamma * gdiv(p, rho));
        const R ca = gdiv(ah_sqrt(sumsq3<R, MASK, 5>(avg)), ah_sqrt(rho));
        const R cax = gdiv(Bx, ah_sqrt(rho));
        const R cs2 = cs * cs, ca2 = ca * ca;
        cf = ah_sqrt(R(0.5) * (cs2 + ca2 + ah_sqrt(sq(cs2 + ca2) - R(4) * cs2 * (cax * cax))));
    } else {
        // signed extreme eigenvalues of the primitive Jacobian (mag_field.py:154-159):
        // the spectrum is {0, u, u+-c_a, u+-c_s, u+-c_f}, so max = max(u + c_f, 0)
        const R a2 = gamma * p / rho;
        const R b2 = sumsq3<R, MASK, 5>(avg) / rho;
        const R bn2 = Bx * Bx / rho;
        const R s = a2 + b2;
        const R disc = s * s - R(4) * a2 * bn2;
        const R cf2 = R(0.5) * (s + ah_sqrt(disc > R(0) ? disc : R(0)));
        cf = ah_sqrt(cf2 > R(0) ? cf2 : R(0));
    }
    ap = nmax(R(0), vx + cf);
    am = -nmin(R(0), vx - cf);
}

// corner EMF (mag_field.py:125-187); one thread per corner (i, j)
template <class R, int MASK, int TS, int TT>
__global__ __launch_bounds__(256) void corner_kernel(MhdArgs a) {
    static_assert(TT == 8 && TS % 8 == 0, "one 8 x 8 block of cells per wave");
    using V = VS<MASK>;
    // cell tiles of the y sweep (s = y, t = x): each wave takes an 8 x 8
    // block, so the x-sweep arrays ([y][x]) and the y-sweep arrays and e3
    // ([x][y]) are both read in 64-byte rows
    const int ntiles = wave_uniform(*a.nttiles_dev[1]);
    const R* Dx = reinterpret_cast<const R*>(a.wD[0]);
    const R* Ux = reinterpret_cast<const R*>(a.wU[0]);
    const R* Dy = reinterpret_cast<const R*>(a.wD[1]);
    const R* Uy = reinterpret_cast<const R*>(a.wU[1]);
    const long long P = a.plane;
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const Tile tl = ld_const(a.ttiles[1] + ti);
        const ProbState& st = a.ps[tl.prob];
        if (!ld_const(&st.active)) continue;
        const ProbDesc pd = ld_const(a.pd + tl.prob);
        const int n0 = pd.n0, n1 = pd.n1;
        const R gamma = R(pd.gamma);
        for (int k = threadIdx.x; k < TS * TT; k += blockDim.x) {
            const int r = k & 63;
            const int i = tl.t0 + (r >> 3), j = tl.s0 + 8 * (k >> 6) + (r & 7);
            if (i >= n0 || j >= n1) continue;
            const int cell = i * n1 + j;
            auto ld = [&](const R* arr, long long o) {
                V8<R> w = zero8<R, MASK>();
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (V::live(k)) w.c[k] = arr[V::slot(k) * P + pd.off + o];
                return w;
            };
            // x-sweep arrays are [t = y][s = x], y-sweep arrays [t = x][s = y]
            const V8<R> D_x = ld(Dx, (long long)j * n0 + i), U_x = ld(Ux, (long long)j * n0 + i);
            const V8<R> D_y = ld(Dy, (long long)i * n1 + j), U_y = ld(Uy, (long long)i * n1 + j);
            R spy_p, spy_m, spx_p, spx_m;
            corner_speeds<R, MASK>(ld(Dx, (long long)bcmap(j + 1, n1, pd.bc) * n0 + i), U_x, 1, gamma, a.solver_hll,
                                   spy_p, spy_m);
            corner_speeds<R, MASK>(ld(Dy, (long long)bcmap(i + 1, n0, pd.bc) * n1 + j), U_y, 0, gamma, a.solver_hll,
                                   spx_p, spx_m);
            // roles follow the reference's dict order, which flips every step
            const bool even = st.parity == 0;
            const V8<R>& north = even ? D_y : D_x;
            const V8<R>& south = even ? U_y : U_x;
            const V8<R>& east = even ? D_x : D_y;
            const V8<R>& west = even ? U_x : U_y;
            const R apy = even ? spx_p : spy_p, amy = even ? spx_m : spy_m;
            const R apx = even ? spy_p : spx_p, amx = even ? spy_m : spx_m;
            const R h = R(0.5);
            const R NE = h * (west.c[2] + south.c[2]) * south.c[5] - h * (west.c[1] + south.c[1]) * west.c[6];
            const R NW = h * (east.c[2] + south.c[2]) * south.c[5] - h * (east.c[1] + south.c[1]) * east.c[6];
            const R SE = h * (west.c[2] + north.c[2]) * north.c[5] - h * (west.c[1] + north.c[1]) * west.c[6];
            const R SW = h * (east.c[2] + north.c[2]) * north.c[5] - h * (east.c[1] + north.c[1]) * east.c[6];
            const R e = gdiv(apx * apy * SW + amx * apy * SE + apx * amy * NW + amx * amy * NE, (apx + amx) * (apy + amy)) -
                        gdiv(apy * amy, apy + amy) * (north.c[5] - south.c[5]) +
                        gdiv(apx * amx, apx + amx) * (east.c[6] - west.c[6]);
            a.e3[pd.off + cell] = double(e);
        }
    }
}

// L of Bx, By from EMF differences (evolvers.py:51-58), overriding the
// hydrodynamic flux difference of those components
template <int MASK>
__global__ __launch_bounds__(256) void magL_kernel(MhdArgs a) {
    using V = VS<MASK>;
    const int ntiles = wave_uniform(*a.ntiles_dev);
    const long long P = a.plane;
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const Tile tl = ld_const(a.tiles + ti);
        if (!ld_const(&a.ps[tl.prob].active)) continue;
        const ProbDesc pd = ld_const(a.pd + tl.prob);
        const DivDx bydx(pd.dx);
        const int n0 = pd.n0, n1 = pd.n1;
        for (int c = threadIdx.x; c < tl.t0; c += blockDim.x) {
            const int cell = tl.s0 + c;
            const int i = cell / n1, j = cell % n1;
            const double* e = a.e3 + pd.off;
            const double e0 = e[cell];
            const double dy = e[(long long)i * n1 + bcmap(j + 1, n1, pd.bc)] - e0;
            const double dxd = e[(long long)bcmap(i + 1, n0, pd.bc) * n1 + j] - e0;
            a.L[V::slot(5) * P + pd.off + cell] = -bydx(dy);
            a.L[V::slot(6) * P + pd.off + cell] = bydx(dxd);
        }
    }
}

// stage 0: B_x, B_y of the state := the sweeps' face values (evolvers.py:73-76)
template <class R, int MASK, int TS, int TT>
__global__ __launch_bounds__(256) void bface_kernel(MhdArgs a) {
    static_assert(TT == 8 && TS % 8 == 0, "one 8 x 8 block of cells per wave");
    using V = VS<MASK>;
    // cell tiles of the y sweep, 8 x 8 cells per wave (Fx is [y][x], U and Fy [x][y])
    const int ntiles = wave_uniform(*a.nttiles_dev[1]);
    const long long P = a.plane;
    const R* Fx = reinterpret_cast<const R*>(a.src[0]);
    const R* Fy = reinterpret_cast<const R*>(a.src[1]);
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const Tile tl = ld_const(a.ttiles[1] + ti);
        const ProbState& st = a.ps[tl.prob];
        if (!ld_const(&st.active)) continue;
        // the state's B is kept so that a step that fails in a later stage
        // can hand back the last good state (copyB_kernel restores it)
        const bool failed = ld_const(&st.fail_flag) != 0;
        const ProbDesc pd = ld_const(a.pd + tl.prob);
        const int n0 = pd.n0, n1 = pd.n1;
        for (int k = threadIdx.x; k < TS * TT; k += blockDim.x) {
            const int r = k & 63;
            const int i = tl.t0 + (r >> 3), j = tl.s0 + 8 * (k >> 6) + (r & 7);
            if (i >= n0 || j >= n1) continue;
            const int cell = i * n1 + j;
            const long long ox = V::slot(5) * P + pd.off + cell, oy = V::slot(6) * P + pd.off + cell;
            a.bsave[pd.off + cell] = a.U[ox];
            a.bsave[P + pd.off + cell] = a.U[oy];
            if (failed) continue;
            a.U[ox] = double(Fx[V::slot(5) * P + pd.off + (long long)j * n0 + i]);
            a.U[oy] = double(Fy[V::slot(6) * P + pd.off + (long long)i * n1 + j]);
        }
    }
}

// inverse reconstruction of B after a stage (mag_field.py:191-210): writes
// the refined Bx, By of register K into two scratch planes.  Per axis (sweep
// frame: s along the axis, t across it): face-centred values along t
// (fv.py:67-85, 'face'), face -> cell centre along s (mag_field.py:203), then
// centre -> average along s and t (fv.py:67-85, 'cntr').  A tile of TT x TS
// cells (x, y) stages both B planes and each intermediate once on the window
// x in [x0-2, x0+TT+3), y in [y0-2, y0+TS+3); a padded value at window
// position p is the one at e(p) (clamped for edge boundaries, the position
// itself for periodic ones, whose window holds the wrapped values).
template <int MASK, int TS, int TT>
__global__ __launch_bounds__(256) void refine_kernel(MhdArgs a) {
    using V = VS<MASK>;
    constexpr int WX = TT + 5, WY = TS + 5;
    __shared__ double Bw[2][WX][WY];
    __shared__ double fcw[2][WX][WY];
    __shared__ double ccw[2][WX][WY];
    const int ntiles = wave_uniform(*a.nttiles_dev[1]);
    const long long P = a.plane;
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const Tile tl = ld_const(a.ttiles[1] + ti);   // y-sweep tiles: s = y, t = x
        const ProbState& st = a.ps[tl.prob];
        if (!ld_const(&st.active) || ld_const(&st.fail_flag)) continue;
        const ProbDesc pd = ld_const(a.pd + tl.prob);
        const int n0 = pd.n0, n1 = pd.n1, bc = pd.bc;
        const int bx = tl.t0 - 2, by = tl.s0 - 2;
        // canonical window index of window index w (e(p) of its position)
        auto cx = [&](int w) { return bc == EDGE ? clampi(bx + w, 0, n0 - 1) - bx : w; };
        auto cy = [&](int w) { return bc == EDGE ? clampi(by + w, 0, n1 - 1) - by : w; };
        const double* B0 = a.K + V::slot(5) * P + pd.off;
        const double* B1 = a.K + V::slot(6) * P + pd.off;
        for (int k = threadIdx.x; k < WX * WY; k += blockDim.x) {
            const int wx = k / WY, wy = k % WY;
            const long long g = (long long)bcmap(bx + wx, n0, bc) * n1 + bcmap(by + wy, n1, bc);
            Bw[0][wx][wy] = B0[g];
            Bw[1][wx][wy] = B1[g];
        }
        __syncthreads();
        // face-centred: axis 0 along y (t), axis 1 along x (t)
        for (int k = threadIdx.x; k < WX * WY; k += blockDim.x) {
            const int wx = k / WY, wy = k % WY;
            if (wy >= 1 && wy < WY - 1) {
                const int qx = cx(wx), qy = cy(wy);
                const double x0 = Bw[0][qx][qy];
                fcw[0][wx][wy] = x0 - K24 * ((Bw[0][qx][cy(qy + 1)] - x0) - (x0 - Bw[0][qx][cy(qy - 1)]));
            }
            if (wx >= 1 && wx < WX - 1) {
                const int qx = cx(wx), qy = cy(wy);
                const double x0 = Bw[1][qx][qy];
                fcw[1][wx][wy] = x0 - K24 * ((Bw[1][cx(qx + 1)][qy] - x0) - (x0 - Bw[1][cx(qx - 1)][qy]));
            }
        }
        __syncthreads();
        // face -> cell centre along s: axis 0 along x, axis 1 along y
        for (int k = threadIdx.x; k < WX * WY; k += blockDim.x) {
            const int wx = k / WY, wy = k % WY;
            if (wx >= 1 && wx < WX - 2 && wy >= 1 && wy < WY - 1) {
                const int qx = cx(wx), qy = cy(wy);
                ccw[0][wx][wy] = -1.0 / 16.0 * (fcw[0][cx(qx - 1)][qy] + fcw[0][cx(qx + 2)][qy]) +
                                 9.0 / 16.0 * (fcw[0][qx][qy] + fcw[0][cx(qx + 1)][qy]);
            }
            if (wy >= 1 && wy < WY - 2 && wx >= 1 && wx < WX - 1) {
                const int qx = cx(wx), qy = cy(wy);
                ccw[1][wx][wy] = -1.0 / 16.0 * (fcw[1][qx][cy(qy - 1)] + fcw[1][qx][cy(qy + 2)]) +
                                 9.0 / 16.0 * (fcw[1][qx][qy] + fcw[1][qx][cy(qy + 1)]);
            }
        }
        __syncthreads();
        // centre -> average: first along s, then along t
        for (int k = threadIdx.x; k < TS * TT; k += blockDim.x) {
            const int ix = k / TS, iy = k % TS;
            const int i = tl.t0 + ix, j = tl.s0 + iy;
            if (i >= n0 || j >= n1) continue;
            const int wx = ix + 2, wy = iy + 2;
            const long long o = pd.off + (long long)i * n1 + j;
            {
                const double c0 = ccw[0][wx][wy];
                double r = c0 + K24 * ((ccw[0][cx(wx + 1)][wy] - c0) - (c0 - ccw[0][cx(wx - 1)][wy]));
                r = r + K24 * ((ccw[0][wx][cy(wy + 1)] - c0) - (c0 - ccw[0][wx][cy(wy - 1)]));
                a.scratch[o] = r;
            }
            {
                const double c0 = ccw[1][wx][wy];
                double r = c0 + K24 * ((ccw[1][wx][cy(wy + 1)] - c0) - (c0 - ccw[1][wx][cy(wy - 1)]));
                r = r + K24 * ((ccw[1][cx(wx + 1)][wy] - c0) - (c0 - ccw[1][cx(wx - 1)][wy]));
                a.scratch[P + o] = r;
            }
        }
        __syncthreads();  // the next tile reuses the windows
    }
}

template <int MASK>
__global__ __launch_bounds__(256) void copyB_kernel(MhdArgs a, double* dst) {
    using V = VS<MASK>;
    const int ntiles = wave_uniform(*a.ntiles_dev);
    const long long P = a.plane;
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const Tile tl = ld_const(a.tiles + ti);
        const ProbState& st = a.ps[tl.prob];
        if (!ld_const(&st.active)) continue;
        const bool failed = ld_const(&st.fail_flag) != 0;
        const ProbDesc pd = ld_const(a.pd + tl.prob);
        for (int c = threadIdx.x; c < tl.t0; c += blockDim.x) {
            const long long o = pd.off + tl.s0 + c;
            if (failed) {
                // the step failed: the state keeps the B it had when the step began
                a.U[V::slot(5) * P + o] = a.bsave[o];
                a.U[V::slot(6) * P + o] = a.bsave[P + o];
            } else {
                dst[V::slot(5) * P + o] = a.scratch[o];
                dst[V::slot(6) * P + o] = a.scratch[P + o];
            }
        }
    }
}

}  // namespace ah

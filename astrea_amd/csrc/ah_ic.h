// Initial states built on the device (functions/constructor.py:11-109 for
// piecewise-constant problems): primitive data at the cell centres of
// [0,1]^d, converted to conservative cell averages as constructor.py:105-109
// does -- the 4th-order conversion for PPM/WENO (generic.py:250-255), then
// the h^2/24 Laplacian for every scheme.  Each elementwise operation is the
// one the numpy pipeline performs, in its order, so the result is that
// pipeline's to the last bit (FP64, no contraction in this unit).
#pragma once
#include "ah_kernels.h"

namespace ah {

struct IcArgs {
    double* state;      // bound state planes
    long long plane;    // plane stride (cells of the plan)
    long long off;      // first cell of the problem
    int n0, n1, dims, bc;
    int nv;
    int varidx[8];      // component stored in plane v
    int kind;           // 0: 1D Riemann split at x0; 1: 2D quadrants split at (x0, y0)
    int high_order;
    double gamma, x0, y0;
    double w[4][8];     // primitive state per region
};

// numpy.linspace(-dx/2, 1 + dx/2, n + 2)[i + 1] with dx = 1/n
// (constructor.py:13-16): arange(num) * step + start, step = delta / (num - 1)
__device__ inline double ic_centre(int i, int n) {
    const double dx = 1.0 / n;
    const double st = 0.0 - dx / 2, sp = 1.0 + dx / 2;
    const double step = (sp - st) / (double)(n + 1);
    return (double)(i + 1) * step + st;
}

__device__ inline void ic_point(const IcArgs& a, int i, int j, double* w) {
    int r;
    const double x = ic_centre(i, a.n0);
    if (a.kind == 0) {
        r = x <= a.x0 ? 0 : 1;
    } else {
        const double y = ic_centre(j, a.n1);
        const bool left = x <= a.x0, low = y < a.y0;
        r = low ? (left ? 0 : 1) : (left ? 2 : 3);
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) w[c] = a.w[r][c];
}

// sqrt(v0^2 + v1^2 + v2^2)^2 as the numpy norm round trip
__device__ inline double ic_norm2(double v0, double v1, double v2) {
    const double n = ::sqrt((v0 * v0 + v1 * v1) + v2 * v2);
    return n * n;
}

__device__ inline void ic_p2c(const double* w, double g, double* q) {
#pragma unroll
    for (int c = 0; c < 8; ++c) q[c] = w[c];
    q[4] = w[4] / (g - 1.0) + 0.5 * (w[0] * ic_norm2(w[1], w[2], w[3]) + ic_norm2(w[5], w[6], w[7]));
    q[1] = w[1] * w[0];
    q[2] = w[2] * w[0];
    q[3] = w[3] * w[0];
}

__device__ inline int ic_map(int p, int n, int bc) {
    return bc == WRAP ? ((p % n) + n) % n : clampi(p, 0, n - 1);
}

__global__ __launch_bounds__(256) void ic_kernel(IcArgs a) {
    const long long cells = (long long)a.n0 * a.n1;
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells) return;
    const int i = (int)(cell / a.n1), j = (int)(cell % a.n1);
    constexpr double K24v = 1.0 / 24.0;
    // neighbour (i, j) shifted by +-1 along axis ax, boundary-mapped
    auto nb = [&](int ii, int jj, int ax, int d, int& oi, int& oj) {
        oi = ax == 0 ? ic_map(ii + d, a.n0, a.bc) : ii;
        oj = ax == 1 ? ic_map(jj + d, a.n1, a.bc) : jj;
    };
    // conservative state before the final Laplacian, at (ii, jj)
    auto qat = [&](int ii, int jj, double* q) {
        double w[8];
        ic_point(a, ii, jj, w);
        if (!a.high_order) {
            ic_p2c(w, a.gamma, q);
            return;
        }
        double ww[8], acc[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            ww[c] = w[c];
            acc[c] = 0.0;
        }
        double pw[8];
        ic_p2c(w, a.gamma, pw);
        for (int ax = 0; ax < a.dims; ++ax) {
            int ui, uj, li, lj;
            nb(ii, jj, ax, 1, ui, uj);
            nb(ii, jj, ax, -1, li, lj);
            double wu[8], wl[8], pu[8], pl[8];
            ic_point(a, ui, uj, wu);
            ic_point(a, li, lj, wl);
            ic_p2c(wu, a.gamma, pu);
            ic_p2c(wl, a.gamma, pl);
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                ww[c] = ww[c] - K24v * ((wu[c] - w[c]) - (w[c] - wl[c]));
                acc[c] = acc[c] + K24v * ((pu[c] - pw[c]) - (pw[c] - pl[c]));
            }
        }
        double qw[8];
        ic_p2c(ww, a.gamma, qw);
#pragma unroll
        for (int c = 0; c < 8; ++c) q[c] = qw[c] + acc[c];
    };
    double q[8], out[8];
    qat(i, j, q);
#pragma unroll
    for (int c = 0; c < 8; ++c) out[c] = q[c];
    for (int ax = 0; ax < a.dims; ++ax) {
        int ui, uj, li, lj;
        nb(i, j, ax, 1, ui, uj);
        nb(i, j, ax, -1, li, lj);
        double qu[8], ql[8];
        qat(ui, uj, qu);
        qat(li, lj, ql);
#pragma unroll
        for (int c = 0; c < 8; ++c) out[c] = out[c] + K24v * ((qu[c] - q[c]) - (q[c] - ql[c]));
    }
    for (int v = 0; v < a.nv; ++v) a.state[v * a.plane + a.off + cell] = out[a.varidx[v]];
}

}  // namespace ah

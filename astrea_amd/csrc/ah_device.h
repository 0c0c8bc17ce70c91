// Device building blocks of the astrea finite-volume update on gfx950.
//
// Everything here is templated on the arithmetic type R (double for the plain
// path, float for the split path) and on the live-component mask MASK (bit c:
// component c of rho, v|m x,y,z, p|E, B x,y,z is carried).  Components outside
// the mask are zero in the reference for the problem class the mask names;
// their terms are dropped at compile time.  Dropping a zero term changes a
// result only in the sign of a zero, so FP64 values equal the reference's
// numpy evaluation, which writes every expression in the order used here.
//
// Reference formulas: functions/fv.py, functions/constructor.py,
// num_methods/limiters.py, schemes/*.py, num_methods/solvers.py.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ah {

enum { EDGE = 0, WRAP = 1 };
enum { PCM = 0, PLM = 1, PPM = 2, WENO3 = 3, WENO5 = 4, WENO7 = 5 };
enum { LLF = 0, HLLC = 1, HLLD = 2, DOTS = 3, ES = 4, LW = 5 };
enum { MINMOD = 0, VANLEER = 1, OSPRE = 2, VANALBADA1 = 3, KOREN = 4, SUPERBEE = 5 };
enum { FLAT_OFF = 0, FLAT_ASTREA = 1, FLAT_COLELLA90 = 2 };

// Python float literals of the reference, rounded once to double.
constexpr double K24 = 1.0 / 24.0;
constexpr double K7_12 = 7.0 / 12.0;
constexpr double K1_12 = 1.0 / 12.0;

template <int MASK>
struct VS {
    static constexpr __host__ __device__ bool live(int c) { return (MASK >> c) & 1; }
    static constexpr __host__ __device__ int slot(int c) {
        int k = 0;
        for (int i = 0; i < c; ++i) k += (MASK >> i) & 1;
        return k;
    }
    static constexpr int NV = ((MASK >> 0) & 1) + ((MASK >> 1) & 1) + ((MASK >> 2) & 1) + ((MASK >> 3) & 1) +
                              ((MASK >> 4) & 1) + ((MASK >> 5) & 1) + ((MASK >> 6) & 1) + ((MASK >> 7) & 1);
    static constexpr bool MHD = ((MASK >> 5) & 7) != 0;
    // component index of storage slot s
    static constexpr __host__ __device__ int comp(int s) {
        int k = 0;
        for (int c = 0; c < 8; ++c)
            if ((MASK >> c) & 1) {
                if (k == s) return c;
                ++k;
            }
        return -1;
    }
};

template <class R>
struct V8 {
    R c[8];
};

template <class R, int MASK>
__device__ inline V8<R> zero8() {
    V8<R> z;
#pragma unroll
    for (int i = 0; i < 8; ++i) z.c[i] = R(0);
    return z;
}

// precise (IEEE) square root and absolute value for each arithmetic type
__device__ inline double ah_sqrt(double x) { return ::sqrt(x); }
__device__ inline float ah_sqrt(float x) { return ::sqrtf(x); }
__device__ inline double ah_fabs(double x) { return ::fabs(x); }
__device__ inline float ah_fabs(float x) { return ::fabsf(x); }
__device__ inline bool ah_isfinite(double x) { return __builtin_isfinite(x); }
__device__ inline bool ah_isfinite(float x) { return __builtin_isfinite(x); }

// ---- elementary operations (functions/fv.py:19-38) --------------------------
template <class R>
__device__ inline R gdiv(R a, R b) { return b != R(0) ? a / b : R(0); }   // fv.divide
template <class R>
__device__ inline R nmin(R a, R b) { return (a < b || a != a) ? a : b; }  // np.minimum (NaN-propagating)
template <class R>
__device__ inline R nmax(R a, R b) { return (a > b || a != a) ? a : b; }  // np.maximum
template <class R>
__device__ inline R nsign(R x) { return x > R(0) ? R(1) : (x < R(0) ? R(-1) : x); }
// x / dx of the flux differences.  When dx is a power of two whose
// reciprocal is representable, x * (1/dx) and x / dx are the correctly
// rounded values of the same real number, hence equal for every x (the
// division is replaced by a multiplication); otherwise it divides.
struct DivDx {
    double dx, inv;
    bool pow2;
    __device__ explicit DivDx(double d) : dx(d) {
        int e = 0;
        const double m = frexp(d, &e);  // d = m 2^e, m in [0.5, 1)
        pow2 = m == 0.5 && 1 - e <= 1023;
        inv = pow2 ? ldexp(1.0, 1 - e) : 0.0;
    }
    template <class R>
    __device__ R operator()(R x) const { return pow2 ? x * R(inv) : x / R(dx); }
};

// Hardware min/max (v_min_f64 / v_max_f64).  On numbers they equal
// np.minimum / np.maximum except in the sign of a zero result; given a NaN
// they return the other operand where numpy returns NaN.  Used only where the
// operands hold no negative zero (absolute values) or the sign of a zero
// result cannot change what is computed from it, and where a NaN operand can
// only come from a NaN cell of the stencil, whose own interface states are
// NaN through the arithmetic, so that the step fails as numpy's does.
__device__ inline double hmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ inline float hmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ inline double hmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ inline float hmax(float a, float b) { return __builtin_fmaxf(a, b); }
template <class R>
__device__ inline R sq(R x) { return x * x; }

// ((x0^2 + x1^2) + x2^2) over components B0..B0+2, dead ones skipped
template <class R, int MASK, int B0>
__device__ inline R sumsq3(const V8<R>& w) {
    using V = VS<MASK>;
    R acc = R(0);
    bool first = true;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (V::live(B0 + i)) {
            R t = w.c[B0 + i] * w.c[B0 + i];
            acc = first ? t : acc + t;
            first = false;
        }
    return acc;
}
template <class R, int MASK, int B0>
constexpr int nlive3() { return VS<MASK>::live(B0) + VS<MASK>::live(B0 + 1) + VS<MASK>::live(B0 + 2); }

// fv.norm(x)**2: the square of the rounded norm.  With one live component
// ah_sqrt(fl(x*x)) == |x| exactly, so the square is fl(x*x).
template <class R, int MASK, int B0>
__device__ inline R norm2(const V8<R>& w) {
    if constexpr (nlive3<R, MASK, B0>() == 0) return R(0);
    else if constexpr (nlive3<R, MASK, B0>() == 1) return sumsq3<R, MASK, B0>(w);
    else {
        R n = ah_sqrt(sumsq3<R, MASK, B0>(w));
        return n * n;
    }
}
// np.sum(v*B, axis=-1) over the velocity / field components that are both live
template <class R, int MASK>
__device__ inline R dot_vB(const V8<R>& w) {
    using V = VS<MASK>;
    R acc = R(0);
    bool first = true;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (V::live(1 + i) && V::live(5 + i)) {
            R t = w.c[1 + i] * w.c[5 + i];
            acc = first ? t : acc + t;
            first = false;
        }
    return acc;
}

// ---- variable conversions (fv.py:49-54, 89-101) ------------------------------
template <class R, int MASK>
__device__ inline V8<R> prim_to_cons(const V8<R>& w, R gamma) {
    using V = VS<MASK>;
    V8<R> q = w;
    R e = w.c[4] / (gamma - R(1));
    if constexpr (V::MHD)
        q.c[4] = e + R(0.5) * (w.c[0] * norm2<R, MASK, 1>(w) + norm2<R, MASK, 5>(w));
    else
        q.c[4] = e + R(0.5) * (w.c[0] * norm2<R, MASK, 1>(w));
#pragma unroll
    for (int i = 1; i < 4; ++i)
        if (V::live(i)) q.c[i] = w.c[i] * w.c[0];
    return q;
}

template <class R, int MASK>
__device__ inline V8<R> cons_to_prim(const V8<R>& q, R gamma) {
    using V = VS<MASK>;
    V8<R> w = q;
#pragma unroll
    for (int i = 1; i < 4; ++i)
        if (V::live(i)) w.c[i] = gdiv(q.c[i], q.c[0]);
    if constexpr (V::MHD)
        w.c[4] = (gamma - R(1)) * (q.c[4] - R(0.5) * (q.c[0] * norm2<R, MASK, 1>(w) + norm2<R, MASK, 5>(q)));
    else
        w.c[4] = (gamma - R(1)) * (q.c[4] - R(0.5) * (q.c[0] * norm2<R, MASK, 1>(w)));
    return w;
}

// ---- physical flux along axis AX (constructor.py:113-125) --------------------
template <class R, int MASK, int AX>
__device__ inline V8<R> phys_flux(const V8<R>& w, R gamma) {
    using V = VS<MASK>;
    constexpr int A = AX % 3, O = (AX + 1) % 3, C = (AX + 2) % 3;
    V8<R> f = zero8<R, MASK>();
    const R rho = w.c[0], u = w.c[AX + 1], p = w.c[4];
    const R ru = rho * u;
    f.c[0] = ru;
    if constexpr (V::MHD) {
        const R b2 = norm2<R, MASK, 5>(w);
        const R Ba = V::live(AX + 5) ? w.c[AX + 5] : R(0);
        f.c[A + 1] = rho * (u * u) + p + R(0.5) * b2 - Ba * Ba;
        if (V::live(O + 1)) f.c[O + 1] = ru * w.c[O + 1] - Ba * w.c[O + 5];
        if (V::live(C + 1)) f.c[C + 1] = ru * w.c[C + 1] - Ba * w.c[C + 5];
        f.c[4] = u * (R(0.5) * rho * norm2<R, MASK, 1>(w) + (gamma * p) / (gamma - R(1)) + b2) - Ba * dot_vB<R, MASK>(w);
        if (V::live(O + 5)) f.c[O + 5] = w.c[O + 5] * u - Ba * w.c[O + 1];
        if (V::live(C + 5)) f.c[C + 5] = w.c[C + 5] * u - Ba * w.c[C + 1];
    } else {
        f.c[A + 1] = rho * (u * u) + p;
        if (V::live(O + 1)) f.c[O + 1] = ru * w.c[O + 1];
        if (V::live(C + 1)) f.c[C + 1] = ru * w.c[C + 1];
        f.c[4] = u * (R(0.5) * rho * norm2<R, MASK, 1>(w) + (gamma * p) / (gamma - R(1)));
    }
    return f;
}

// ---- Roe average (constructor.py:167-176) and arithmetic mean (plm.py:84) ----
template <class R, int MASK>
__device__ inline V8<R> roe_avg(const V8<R>& wp, const V8<R>& wm) {
    using V = VS<MASK>;
    V8<R> a = zero8<R, MASK>();
    const R rm = ah_sqrt(wm.c[0]), rp = ah_sqrt(wp.c[0]);
    const R s = rm + rp;
    a.c[0] = rm * rp;
#pragma unroll
    for (int i = 1; i < 4; ++i)
        if (V::live(i)) a.c[i] = gdiv(wp.c[i] * rp + wm.c[i] * rm, s);
    a.c[4] = gdiv(rp * wp.c[4] + rm * wm.c[4], s);
#pragma unroll
    for (int i = 5; i < 8; ++i)
        if (V::live(i)) a.c[i] = gdiv(wp.c[i] * rm + wm.c[i] * rp, s);
    return a;
}
template <class R, int MASK>
__device__ inline V8<R> arith_avg(const V8<R>& wp, const V8<R>& wm) {
    V8<R> a = zero8<R, MASK>();
#pragma unroll
    for (int i = 0; i < 8; ++i)
        if (VS<MASK>::live(i)) a.c[i] = R(0.5) * (wp.c[i] + wm.c[i]);
    return a;
}

// ---- spectral radius of the primitive Jacobian (replaces fv.py:157-161) ------
// The B-normal row of constructor.make_Jacobian is zero, so its spectrum is
// {0} plus that of the 7x7 primitive MHD Jacobian: u, u+-c_a, u+-c_s, u+-c_f.
// max|lambda| = |u| + c_f.  A non-finite state returns NaN (the reference's
// LAPACK call raises on it), which the caller folds into the failure flag.
template <class R, int MASK, int AX>
__device__ inline R spectral_radius(const V8<R>& w, R gamma) {
    using V = VS<MASK>;
    bool fin = true;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        if (V::live(i)) fin = fin && ah_isfinite(w.c[i]);
    const R rho = w.c[0], u = w.c[AX + 1], p = w.c[4];
    const R a2 = gamma * p / rho;
    R r;
    if constexpr (!V::MHD) {
        r = a2 >= R(0) ? ah_fabs(u) + ah_sqrt(a2) : ah_sqrt(ah_fabs(u * u - a2));
    } else {
        const R b2 = sumsq3<R, MASK, 5>(w) / rho;
        const R Bn = V::live(AX + 5) ? w.c[AX + 5] : R(0);
        const R bn2 = Bn * Bn / rho;
        const R s = a2 + b2;
        const R disc = s * s - R(4) * a2 * bn2;
        const R cf2 = R(0.5) * (s + ah_sqrt(disc > R(0) ? disc : R(0)));
        r = ah_fabs(u) + ah_sqrt(cf2 > R(0) ? cf2 : R(0));
    }
    return fin ? r : R(NAN);
}

// ---- slope limiters (limiters.py:10-49); a backward, b forward difference ----
template <class R>
__device__ inline R slope_limit(int lim, R a, R b) {
    if (lim == MINMOD) {
        const bool ab = a * b > R(0);
        if (ah_fabs(a) < ah_fabs(b) && ab) return a;
        if (ah_fabs(a) >= ah_fabs(b) && ab) return b;
        return R(0);
    }
    const R r = gdiv(a, b);
    R phi;
    switch (lim) {
        case VANLEER: phi = (r + ah_fabs(r)) / (R(1) + ah_fabs(r)); break;
        case OSPRE: phi = R(1.5) * ((r * r + r) / (r * r + r + R(1))); break;
        case VANALBADA1: phi = (r * r + r) / (r * r + R(1)); break;
        case KOREN: phi = nmax(R(0), nmin(nmin(R(2) * r, (R(2) + r) / R(3)), R(2))); break;
        default: phi = nmax(R(0), nmax(nmin(R(2) * r, R(1)), nmin(r, R(2)))); break;
    }
    return phi * b;
}

// ---- PPM 'mc' extrapolant limiter, one variable (limiters.py:82-140) --------
// The reference's grid-global cell_extrema.any() test (limiters.py:102) is
// always true (the first cell with edge ghosts, or any periodic line, has a
// cell extremum), so only that branch exists here.
// Inputs: wS, neighbours wm1/wp1/wm2/wp2, face values fL/fR, curvatures
// d2 at i-1, i, i+1 and third differences d3 at i-2..i+2.
template <class R>
__device__ inline void mc_limit(R wS, R wm1, R wp1, R wm2, R wp2, R fL, R fR, R d2m, R d2c, R d2p, R d3m2, R d3m1,
                                R d30, R d3p2, R& wL, R& wR) {
    const R C = R(1.25);
    const R dwm = wS - fL, dwp = fR - wS;
    wL = fL;
    wR = fR;
    const R d2w = R(6) * (fL - R(2) * wS + fR);
    const bool cell_extrema = (dwm * dwp <= R(0)) || ((wS - wm2) * (wp2 - wS) <= R(0));
    // min/max over absolute values: no negative zeros (hmin/hmax exact)
    const R lim_curv = nsign(d2c) * hmin(hmin(ah_fabs(d2w), C * ah_fabs(d2c)), hmin(C * ah_fabs(d2p), C * ah_fabs(d2m)));
    const R d2w_lim = cell_extrema ? lim_curv : R(0);
    const bool sens = ah_fabs(d2w) > R(1e-12) * hmax(ah_fabs(wS), hmax(hmax(ah_fabs(wm1), ah_fabs(wp1)), hmax(ah_fabs(wm2), ah_fabs(wp2))));
    const R rho = sens ? gdiv(d2w_lim, d2w) : R(0);
    // d3min/d3max enter only through |.| and d3max - d3min, which compare
    // the same against 0.1 max(|d3max|, |d3min|) whatever the zeros' signs
    const R d3min = hmin(hmin(d3m1, d30), hmin(d3m2, d3p2));
    const R d3max = hmax(hmax(d3m1, d30), hmax(d3m2, d3p2));
    const bool roundoff = (rho < R(1.0 - 1e-12)) || (R(0.1) * hmax(ah_fabs(d3max), ah_fabs(d3min)) <= (d3max - d3min));
    if ((dwm * dwp < R(0)) && roundoff) {
        wL = wS - rho * dwm;
        wR = wS + rho * dwp;
    }
    if (roundoff && ah_fabs(dwm) >= R(2) * ah_fabs(dwp)) wL = wS - R(2) * (R(1) - rho) * dwp - rho * dwm;
    if (roundoff && ah_fabs(dwp) >= R(2) * ah_fabs(dwm)) wR = wS + R(2) * (R(1) - rho) * dwm + rho * dwp;
}

// ---- WENO (weno.py:22-149), component-wise, eps = 1e-6 -----------------------
template <class R>
__device__ inline void weno3(R m1, R z, R p1, R& wL, R& wR) {
    const R eps = R(1e-6);
    const R b0 = sq(z - m1), b1 = sq(p1 - z);
    const R g0 = R(1.0 / 3.0), g1 = R(2.0 / 3.0);
    const R e0 = sq(b0 + eps), e1 = sq(b1 + eps);
    const R a0g0 = g0 / e0, a1g1 = g1 / e1, a1g0 = g0 / e1, a0g1 = g1 / e0;
    wR = (a0g0 / (a0g0 + a1g1)) * (R(1.5) * z - R(0.5) * m1) + (a1g1 / (a0g0 + a1g1)) * (R(0.5) * z + R(0.5) * p1);
    wL = (a1g0 / (a0g1 + a1g0)) * (R(1.5) * z - R(0.5) * p1) + (a0g1 / (a0g1 + a1g0)) * (R(0.5) * z + R(0.5) * m1);
}

template <class R>
__device__ inline void weno5(R m2, R m1, R z, R p1, R p2, R& wL, R& wR) {
    const R eps = R(1e-6);
    const R c13 = R(13.0 / 12.0), c14 = R(1.0 / 4.0);
    const R b0 = c13 * sq(m2 - R(2) * m1 + z) + c14 * sq(m2 - R(4) * m1 + R(3) * z);
    const R b1 = c13 * sq(m1 - R(2) * z + p1) + c14 * sq(m1 - p1);
    const R b2 = c13 * sq(z - R(2) * p1 + p2) + c14 * sq(R(3) * z - R(4) * p1 + p2);
    const R g0 = R(1.0 / 10.0), g1 = R(3.0 / 5.0), g2 = R(3.0 / 10.0);
    const R e0 = sq(b0 + eps), e1 = sq(b1 + eps), e2 = sq(b2 + eps);
    const R A0g0 = g0 / e0, A1g1 = g1 / e1, A2g2 = g2 / e2, A0g2 = g2 / e0, A2g0 = g0 / e2;
    const R t3 = R(1.0 / 3.0), t76 = R(7.0 / 6.0), t116 = R(11.0 / 6.0), m16 = R(-1.0 / 6.0), t56 = R(5.0 / 6.0);
    const R sR = A0g0 + A1g1 + A2g2;
    wR = (A0g0 / sR) * (t3 * m2 - t76 * m1 + t116 * z) + (A1g1 / sR) * (m16 * m1 + t56 * z + t3 * p1) +
         (A2g2 / sR) * (t3 * z + t56 * p1 - R(1.0 / 6.0) * p2);
    const R sL = A0g2 + A1g1 + A2g0;
    wL = (A0g2 / sL) * (t3 * z + t56 * m1 - R(1.0 / 6.0) * m2) + (A1g1 / sL) * (m16 * p1 + t56 * z + t3 * m1) +
         (A2g0 / sL) * (t3 * p2 - t76 * p1 + t116 * z);
}

template <class R>
__device__ inline void weno7(R m3, R m2, R m1, R z, R p1, R p2, R p3, R& wL, R& wR) {
    const R eps = R(1e-6);
    const R b0 = m3 * (R(547) * m3 - R(3882) * m2 + R(4642) * m1 - R(1854) * z) +
                 m2 * (R(7043) * m2 - R(17246) * m1 + R(7042) * z) + m1 * (R(11003) * m1 - R(9402) * z) + z * (R(2107) * z);
    const R b1 = m2 * (R(267) * m2 - R(1642) * m1 + R(1602) * z - R(494) * p1) +
                 m1 * (R(2843) * m1 - R(5966) * z + R(1922) * p1) + z * (R(3443) * z - R(2522) * p1) + p1 * (R(547) * p1);
    const R b2 = m1 * (R(547) * m1 - R(2522) * z + R(1922) * p1 - R(494) * p2) +
                 z * (R(3443) * z - R(5966) * p1 + R(1602) * p2) + p1 * (R(2843) * p1 - R(1642) * p2) + p2 * (R(267) * p2);
    const R b3 = z * (R(2107) * z - R(9402) * p1 + R(7042) * p2 - R(1854) * p3) +
                 p1 * (R(11003) * p1 - R(17246) * p2 + R(4642) * p3) + p2 * (R(7043) * p2 - R(3882) * p3) + p3 * (R(547) * p3);
    const R g0 = R(1.0 / 35.0), g1 = R(12.0 / 35.0), g2 = R(18.0 / 35.0), g3 = R(4.0 / 35.0);
    const R e0 = sq(b0 + eps), e1 = sq(b1 + eps), e2 = sq(b2 + eps), e3 = sq(b3 + eps);
    const R q14 = R(1.0 / 4.0), q1312 = R(13.0 / 12.0), q2312 = R(23.0 / 12.0), q2512 = R(25.0 / 12.0),
            q112 = R(1.0 / 12.0), q512 = R(5.0 / 12.0), q712 = R(7.0 / 12.0);
    {
        const R A0 = g0 / e0, A1 = g1 / e1, A2 = g2 / e2, A3 = g3 / e3;
        const R s = A0 + A1 + A2 + A3;
        wR = (A0 / s) * (R(-1.0 / 4.0) * m3 + q1312 * m2 - q2312 * m1 + q2512 * z) +
             (A1 / s) * (q112 * m2 - q512 * m1 + q1312 * z + q14 * p1) +
             (A2 / s) * (R(-1.0 / 12.0) * m1 + q712 * z + q712 * p1 - q112 * p2) +
             (A3 / s) * (q14 * z + q1312 * p1 - q512 * p2 + q112 * p3);
    }
    {
        const R A0 = g3 / e0, A1 = g2 / e1, A2 = g1 / e2, A3 = g0 / e3;
        const R s = A0 + A1 + A2 + A3;
        wL = (A0 / s) * (q14 * z + q1312 * m1 - q512 * m2 + q112 * m3) +
             (A1 / s) * (R(-1.0 / 12.0) * p1 + q712 * z + q712 * m1 - q112 * m2) +
             (A2 / s) * (q112 * p2 - q512 * p1 + q1312 * z + q14 * m1) +
             (A3 / s) * (R(-1.0 / 4.0) * p3 + q1312 * p2 - q2312 * p1 + q2512 * z);
    }
}

// ---- PPM slope flattener (schemes/ppm.py:111-134, Colella 1990) -------------
// chi~ of one cell from pressure p(i-2..i+2) and normal velocity u(i-1), u(i+1).
template <class R>
__device__ inline R flat_chi(int variant, R pm2, R pm1, R pp1, R pp2, R um1, R up1) {
    const R delta = R(0.33), z0 = R(0.75), z1 = R(0.85);
    const R dp1 = pp1 - pm1;
    const R z = gdiv(ah_fabs(dp1), ah_fabs(pp2 - pm2));
    R chi = R(1) - gdiv(z - z0, R(0.85 - 0.75));
    if (z > z1) chi = R(0);
    if (z < z0) chi = R(1);
    const bool weak = gdiv(ah_fabs(dp1), nmin(pp1, pm1)) <= delta;
    const bool expanding = (um1 - up1) <= R(0);
    if (variant == FLAT_ASTREA) {
        if (expanding && weak) chi = R(0);
    } else {
        if (expanding || weak) chi = R(1);
    }
    return chi;
}

// ---- Riemann solvers (solvers.py) --------------------------------------------
// Interface quantities: w/q/f for the minus (left) and plus (right) states.
template <class R>
struct Iface {
    V8<R> wm, wp, qm, qp, fm, fp;
};

// Local Lax-Friedrichs (solvers.py:69-75)
template <class R, int MASK>
__device__ inline V8<R> flux_llf(const Iface<R>& s, R lam) {
    V8<R> F = zero8<R, MASK>();
#pragma unroll
    for (int i = 0; i < 8; ++i)
        if (VS<MASK>::live(i)) F.c[i] = R(0.5) * (s.fm.c[i] + s.fp.c[i]) - R(0.5) * ((s.qp.c[i] - s.qm.c[i]) * lam);
    return F;
}

// Lax-Wendroff dissipation coefficient at a Jacobian state (solvers.py:82-85):
// ss^2 / max|lambda| with ss the sound speed sqrt(gamma p / rho), the wave
// speed the reference's comment names, and r = max|lambda| (the spectral
// radius).  Paired (k, k+1) and applied as the LLF flux applies its radii.
template <class R>
__device__ inline R lw_coef(const V8<R>& w, R r, R gamma) {
    const R ss = ah_sqrt(gamma * w.c[4] / w.c[0]);
    return gdiv(ss * ss, r);
}

// HLLC (solvers.py:92-138).  sax = the solver's axis argument (the
// enumeration index of the sweep, solvers.py:34-36); the reference writes the
// star momentum at index 1 whatever the axis (127-128), `mom` selects it.
template <class R, int MASK>
__device__ inline V8<R> flux_hllc(const Iface<R>& s, int sax, R gamma, int mom, bool default_minus) {
    using V = VS<MASK>;
    const R rhoL = s.wm.c[0], pL = s.wm.c[4], rhoR = s.wp.c[0], pR = s.wp.c[4];
    const R uL = sax == 0 ? s.wm.c[1] : (V::live(2) ? s.wm.c[2] : R(0));
    const R uR = sax == 0 ? s.wp.c[1] : (V::live(2) ? s.wp.c[2] : R(0));
    const R cL = ah_sqrt(gamma * gdiv(pL, rhoL)), cR = ah_sqrt(gamma * gdiv(pR, rhoR));
    const R sqL = ah_sqrt(rhoL), sqR = ah_sqrt(rhoR);
    const R ssum = sqL + sqR;
    const R u_hat = gdiv(uL * sqL + uR * sqR, ssum);
    const R c2_hat = gdiv(sqL * (cL * cL) + sqR * (cR * cR), ssum) + R(0.5) * sq(uR - uL) * gdiv(sqL * sqR, ssum * ssum);
    const R sc = ah_sqrt(c2_hat);
    const R sL = nmin(uL - cL, u_hat - sc);
    const R sR = nmax(uR + cR, u_hat + sc);
    const R sM = gdiv(pR - pL + rhoL * uL * (sL - uL) - rhoR * uR * (sR - uR), rhoL * (sL - uL) - rhoR * (sR - uR));
    const R kL = gdiv(sL - uL, sL - sM), kR = gdiv(sR - uR, sR - sM);
    const bool useL = (sL <= R(0)) && (R(0) < sM);
    const bool useR = (sM <= R(0)) && (R(0) <= sR);
    const bool useP = sR < R(0);
    // select first, then evaluate one star flux: F(w+) | F*L | F*R | default.
    // The reference's default is F(w+) (solvers.py:133), which is the
    // downwind flux when all waves move right; `default_minus` gives F(w-).
    if (useP) return s.fp;
    if (!(useL || useR)) return default_minus ? s.fm : s.fp;
    V8<R> F = s.fp;
    const bool left = useL && !useR;
    const V8<R>& Q = left ? s.qm : s.qp;
    const V8<R>& Fb = left ? s.fm : s.fp;
    const R k = left ? kL : kR, rho = left ? rhoL : rhoR, u = left ? uL : uR, p = left ? pL : pR;
    const R sS = left ? sL : sR;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (!V::live(i)) continue;
        R qs = Q.c[i] * k;
        if (i == mom) qs = rho * k * sM;
        if (i == 4) qs = qs + k * (sM - u) * (rho * sM + gdiv(p, sS - u));
        F.c[i] = Fb.c[i] + (qs - Q.c[i]) * sS;
    }
    return F;
}

// A value the whole workgroup shares, read from lane 0 so that the compiler
// keeps it (and control flow on it) scalar: loads from memory the kernel also
// writes are vector loads, which would otherwise make tile loops divergent.
__device__ inline int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Read-only tables (tile lists, problem descriptors, flags that only other
// launches write) read through the constant address space: at a uniform
// address this is a scalar load, which the compiler cannot otherwise prove
// safe inside a loop that also stores to global memory.
template <class T>
__device__ inline T ld_const(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "word-sized tables");
    using CW = const __attribute__((address_space(4))) int;
    CW* q = (CW*)p;
    T out;
    int* o = reinterpret_cast<int*>(&out);
#pragma unroll
    for (int i = 0; i < int(sizeof(T) / 4); ++i) o[i] = q[i];
    return out;
}

// ---- split-precision helpers ---------------------------------------------------
// Values (states, interface states, fluxes) stay FP64; only coefficients run in
// FP32.  These variants are not bitwise reference-equal (they use
// E = p * 1/(gamma-1), squared norms without the sqrt/square round trip, and
// refined FP32 reciprocals) but are accurate to ~1e-14 relative.
__device__ inline double rcp_refined(double x) {
    const float r0 = 1.0f / float(x);
    double r = double(r0);
    if (!(r0 != 0.0f) || !ah_isfinite(r0)) return x != 0.0 ? 1.0 / x : 0.0;
    r = r + r * (1.0 - x * r);  // one Newton step: relative error ~ (2^-24)^2
    return r;
}

template <int MASK>
__device__ inline V8<double> cons_to_prim_split(const V8<double>& q, double gamma) {
    using V = VS<MASK>;
    V8<double> w = q;
    const double r = q.c[0] != 0.0 ? rcp_refined(q.c[0]) : 0.0;
#pragma unroll
    for (int i = 1; i < 4; ++i)
        if (V::live(i)) w.c[i] = q.c[i] * r;
    double kin = q.c[0] * sumsq3<double, MASK, 1>(w);
    if constexpr (V::MHD) kin = kin + sumsq3<double, MASK, 5>(q);
    w.c[4] = (gamma - 1.0) * (q.c[4] - 0.5 * kin);
    return w;
}

template <int MASK>
__device__ inline V8<double> prim_to_cons_split(const V8<double>& w, double gamma) {
    using V = VS<MASK>;
    V8<double> q = w;
    double kin = w.c[0] * sumsq3<double, MASK, 1>(w);
    if constexpr (V::MHD) kin = kin + sumsq3<double, MASK, 5>(w);
    q.c[4] = w.c[4] * (1.0 / (gamma - 1.0)) + 0.5 * kin;
#pragma unroll
    for (int i = 1; i < 4; ++i)
        if (V::live(i)) q.c[i] = w.c[i] * w.c[0];
    return q;
}

template <int MASK, int AX>
__device__ inline V8<double> phys_flux_split(const V8<double>& w, double gamma) {
    using V = VS<MASK>;
    constexpr int A = AX % 3, O = (AX + 1) % 3, C = (AX + 2) % 3;
    V8<double> f = zero8<double, MASK>();
    const double rho = w.c[0], u = w.c[AX + 1], p = w.c[4];
    const double ru = rho * u;
    const double g1 = gamma / (gamma - 1.0);
    f.c[0] = ru;
    f.c[A + 1] = ru * u + p;
    if (V::live(O + 1)) f.c[O + 1] = ru * w.c[O + 1];
    if (V::live(C + 1)) f.c[C + 1] = ru * w.c[C + 1];
    f.c[4] = u * (0.5 * rho * sumsq3<double, MASK, 1>(w) + g1 * p);
    if constexpr (V::MHD) {
        const double b2 = sumsq3<double, MASK, 5>(w);
        const double Ba = w.c[AX + 5];
        f.c[A + 1] = f.c[A + 1] + 0.5 * b2 - Ba * Ba;
        if (V::live(O + 1)) f.c[O + 1] = f.c[O + 1] - Ba * w.c[O + 5];
        if (V::live(C + 1)) f.c[C + 1] = f.c[C + 1] - Ba * w.c[C + 5];
        f.c[4] = f.c[4] + u * b2 - Ba * dot_vB<double, MASK>(w);
        if (V::live(O + 5)) f.c[O + 5] = w.c[O + 5] * u - Ba * w.c[O + 1];
        if (V::live(C + 5)) f.c[C + 5] = w.c[C + 5] * u - Ba * w.c[C + 1];
    }
    return f;
}

// HLLC in perturbation form: wave speeds and star factors in FP32, the flux
// assembled in FP64 from F +/- s (Q* - Q) with Q* - Q written through
// (k - 1) = (sM - u) / (sS - sM) so no O(1) cancellation happens in FP32.
template <int MASK>
__device__ inline V8<double> flux_hllc_split(const Iface<double>& s, int sax, double gamma, int mom,
                                             bool default_minus) {
    using V = VS<MASK>;
    const double rhoLd = s.wm.c[0], pLd = s.wm.c[4], rhoRd = s.wp.c[0], pRd = s.wp.c[4];
    const double uLd = sax == 0 ? s.wm.c[1] : (V::live(2) ? s.wm.c[2] : 0.0);
    const double uRd = sax == 0 ? s.wp.c[1] : (V::live(2) ? s.wp.c[2] : 0.0);
    const float rhoL = float(rhoLd), rhoR = float(rhoRd), uL = float(uLd), uR = float(uRd);
    const float pL = float(pLd), pR = float(pRd), g = float(gamma);
    // jumps taken in FP64 before rounding: every speed below is carried as an
    // offset from uL or uR, so sM - u has no cancellation in FP32
    const float du = float(uRd - uLd), dp = float(pRd - pLd);
    const float cL = ah_sqrt(g * gdiv(pL, rhoL)), cR = ah_sqrt(g * gdiv(pR, rhoR));
    const float sqL = ah_sqrt(rhoL), sqR = ah_sqrt(rhoR), ssum = sqL + sqR;
    const float uhL = gdiv(du * sqR, ssum), uhR = -gdiv(du * sqL, ssum);  // u_hat - uL, u_hat - uR
    const float c2_hat = gdiv(sqL * (cL * cL) + sqR * (cR * cR), ssum) + 0.5f * (du * du) * gdiv(sqL * sqR, ssum * ssum);
    const float sc = ah_sqrt(c2_hat);
    const float sLu = nmin(-cL, uhL - sc), sRu = nmax(cR, uhR + sc);  // sL - uL, sR - uR
    const float sL = uL + sLu, sR = uR + sRu;
    // sM = (dp + rhoL uL (sL-uL) - rhoR uR (sR-uR)) / (rhoL (sL-uL) - rhoR (sR-uR)), as offsets
    const float den0 = rhoL * sLu - rhoR * sRu;
    const float sMuL = gdiv(dp - rhoR * sRu * du, den0), sMuR = gdiv(dp - rhoL * sLu * du, den0);
    const float sM = uL + sMuL;
    const bool useL = (sL <= 0.f) && (0.f < sM);
    const bool useR = (sM <= 0.f) && (0.f <= sR);
    if (sR < 0.f) return s.fp;
    if (!(useL || useR)) return default_minus ? s.fm : s.fp;
    const bool left = useL && !useR;
    const V8<double>& Q = left ? s.qm : s.qp;
    const V8<double>& Fb = left ? s.fm : s.fp;
    const float sS = left ? sL : sR, rho = left ? rhoL : rhoR, p = left ? pL : pR;
    const float sSu = left ? sLu : sRu, sMu = left ? sMuL : sMuR;  // sS - u, sM - u
    const float den = sSu - sMu;                                    // sS - sM
    // k = (sS - u)/(sS - sM) (0 where the denominator vanishes, fv.divide)
    const float km1 = den != 0.f ? sMu / den : -1.f;
    const float k = den != 0.f ? sSu / den : 0.f;
    const float eadd = k * sMu * (rho * sM + gdiv(p, sSu));
    const double rho_u = left ? rhoLd * uLd : rhoRd * uRd;
    V8<double> F = Fb;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (!V::live(i)) continue;
        double dq = Q.c[i] * double(km1);
        // rho k sM - Q = rho (u + sS (k - 1)) - Q
        if (i == mom) dq = (left ? rhoLd : rhoRd) * double(sS) * double(km1) + (rho_u - Q.c[i]);
        if (i == 4) dq = dq + double(eadd);
        F.c[i] = Fb.c[i] + dq * double(sS);
    }
    return F;
}

// HLLD (solvers.py:142-232), evaluated on all 8 components exactly as the
// reference writes it (dead components enter as zeros).  SAX = solver axis;
// Bn = normal field of the right cell's centred state (solvers.py:168,183).
// FP64 on the split path too: its region selection is discontinuous (both
// Alfven speeds s_M - B_n / sqrt(rho*), the default F(w-)), and FP32 fast
// speeds flipped a region of the Brio-Wu problem within six steps.
template <class R, int MASK, int SAX>
__device__ inline V8<R> flux_hlld(const Iface<R>& s, R Bn, R gamma) {
    constexpr int A = SAX % 3, O = (SAX + 1) % 3, C = (SAX + 2) % 3;
    auto cfast = [&](const V8<R>& w) {
        const R rho = w.c[0], p = w.c[4];
        const R cs = ah_sqrt(gdiv(gamma * p, rho));
        const R ca = gdiv(ah_sqrt((w.c[5] * w.c[5] + w.c[6] * w.c[6]) + w.c[7] * w.c[7]), ah_sqrt(rho));
        const R cax = gdiv(w.c[5], ah_sqrt(rho));
        const R cs2 = cs * cs, ca2 = ca * ca;
        return ah_sqrt(R(0.5) * (cs2 + ca2 + ah_sqrt(sq(cs2 + ca2) - R(4) * cs2 * (cax * cax))));
    };
    auto n2 = [](R a, R b, R c) {
        R n = ah_sqrt((a * a + b * b) + c * c);
        return n * n;
    };
    auto d3 = [](R a0, R b0, R a1, R b1, R a2, R b2) { return (a0 * b0 + a1 * b1) + a2 * b2; };
    const V8<R>&wL = s.wm, &wR = s.wp, &QL = s.qm, &QR = s.qp;
    const R rhoL = wL.c[0], pL = wL.c[4], rhoR = wR.c[0], pR = wR.c[4];
    const R uL = wL.c[SAX + 1], uR = wR.c[SAX + 1];
    const R cfL = cfast(wL), cfR = cfast(wR);
    const R umin = nmin(uL, uR), cmax = nmax(cfL, cfR);
    const R sL = umin - cmax, sR = umin + cmax;
    const R bL2 = n2(wL.c[5], wL.c[6], wL.c[7]), bR2 = n2(wR.c[5], wR.c[6], wR.c[7]);
    const R den = rhoL * (sL - uL) - rhoR * (sR - uR);
    const R sM = gdiv(pR - pL + rhoL * uL * (sL - uL) - rhoR * uR * (sR - uR) + R(0.5) * bR2 - R(0.5) * bL2, den);
    const R rLs = rhoL * gdiv(sL - uL, sL - sM), rRs = rhoR * gdiv(sR - uR, sR - sM);
    const R sLs = sM - gdiv(wL.c[SAX + 5], ah_sqrt(rLs)), sRs = sM - gdiv(wR.c[SAX + 5], ah_sqrt(rRs));
    const R ps = gdiv(rhoL * (pR + R(0.5) * bR2) * (sL - uL) - rhoR * (pL + R(0.5) * bL2) * (sR - uR) +
                          rhoR * rhoL * (sL - uL) * (sR - uR),
                      den);
    const R Bn2 = Bn * Bn;
    const R dL = rhoL * (sL - uL) * (sL - sM) - Bn2, dR = rhoR * (sR - uR) * (sR - sM) - Bn2;
    const R vyLs = wL.c[O + 1] - Bn * wL.c[O + 5] * gdiv(sM - uL, dL);
    const R vyRs = wR.c[O + 1] - Bn * wR.c[O + 5] * gdiv(sM - uR, dR);
    const R vzLs = wL.c[C + 1] - Bn * wL.c[C + 5] * gdiv(sM - uL, dL);
    const R vzRs = wR.c[C + 1] - Bn * wR.c[C + 5] * gdiv(sM - uR, dR);
    const R ByLs = wL.c[O + 5] * gdiv(rhoL * sq(sL - uL) - Bn2, dL);
    const R ByRs = wR.c[O + 5] * gdiv(rhoR * sq(sR - uR) - Bn2, dR);
    const R BzLs = wL.c[C + 5] * gdiv(rhoL * sq(sL - uL) - Bn2, dL);
    const R BzRs = wR.c[C + 5] * gdiv(rhoR * sq(sR - uR) - Bn2, dR);
    V8<R> QLs, QRs;
#pragma unroll
    for (int i = 0; i < 8; ++i) QLs.c[i] = QRs.c[i] = R(0);
    QLs.c[0] = rLs;
    QRs.c[0] = rRs;
    QLs.c[A + 1] = rhoL * sM;
    QRs.c[A + 1] = rhoR * sM;
    QLs.c[O + 1] = rhoL * vyLs;
    QRs.c[O + 1] = rhoR * vyRs;
    QLs.c[C + 1] = rhoL * vzLs;
    QRs.c[C + 1] = rhoR * vzRs;
    QLs.c[A + 5] = QL.c[A + 5];
    QRs.c[A + 5] = QR.c[A + 5];
    QLs.c[O + 5] = ByLs;
    QRs.c[O + 5] = ByRs;
    QLs.c[C + 5] = BzLs;
    QRs.c[C + 5] = BzRs;
    const R vBL = d3(wL.c[1], wL.c[5], wL.c[2], wL.c[6], wL.c[3], wL.c[7]);
    const R vBR = d3(wR.c[1], wR.c[5], wR.c[2], wR.c[6], wR.c[3], wR.c[7]);
    const R mBLs = d3(QLs.c[1], QLs.c[5], QLs.c[2], QLs.c[6], QLs.c[3], QLs.c[7]);
    const R mBRs = d3(QRs.c[1], QRs.c[5], QRs.c[2], QRs.c[6], QRs.c[3], QRs.c[7]);
    QLs.c[4] = gdiv(QL.c[4] * (sL - uL) - uL * (pL + R(0.5) * bL2) + ps * sM + Bn * (vBL - mBLs), sL - sM);
    QRs.c[4] = gdiv(QR.c[4] * (sR - uR) - uR * (pR + R(0.5) * bR2) + ps * sM + Bn * (vBR - mBRs), sR - sM);
    const R sg = nsign(Bn);
    const R srL = ah_sqrt(rLs), srR = ah_sqrt(rRs), srs = srL + srR, srp = ah_sqrt(rRs * rLs);
    const R vyss = gdiv(vyRs * srR + vyLs * srL + sg * (ByLs - ByRs), srs);
    const R vzss = gdiv(vzRs * srR + vzLs * srL + sg * (BzLs - BzRs), srs);
    const R Byss = gdiv(ByLs * srR + ByRs * srL + sg * (vyLs - vyRs) * srp, srs);
    const R Bzss = gdiv(BzLs * srR + BzRs * srL + sg * (vzLs - vzRs) * srp, srs);
    V8<R> QLss, QRss;
#pragma unroll
    for (int i = 0; i < 8; ++i) QLss.c[i] = QRss.c[i] = R(0);
    QLss.c[0] = rLs;
    QRss.c[0] = rRs;
    QLss.c[A + 1] = sM;
    QRss.c[A + 1] = sM;
    QLss.c[O + 1] = vyss;
    QRss.c[O + 1] = vyss;
    QLss.c[C + 1] = vzss;
    QRss.c[C + 1] = vzss;
    QLss.c[A + 5] = QLs.c[A + 5];
    QRss.c[A + 5] = QRs.c[A + 5];
    QLss.c[O + 5] = Byss;
    QRss.c[O + 5] = Byss;
    QLss.c[C + 5] = Bzss;
    QRss.c[C + 5] = Bzss;
    const R mBLss = d3(QLss.c[1], QLss.c[5], QLss.c[2], QLss.c[6], QLss.c[3], QLss.c[7]);
    const R mBRss = d3(QRss.c[1], QRss.c[5], QRss.c[2], QRss.c[6], QRss.c[3], QRss.c[7]);
    QLss.c[4] = QLs.c[4] - srL * sg * (mBLs - mBLss);
    QRss.c[4] = QRs.c[4] - srR * sg * (mBRs - mBRss);
    // region select (solvers.py:226-231), later masks overwrite earlier ones
    int reg = 0;  // 0: F(w-)
    if ((sL <= R(0)) && (R(0) < sLs)) reg = 1;
    if ((sLs <= R(0)) && (R(0) < sM)) reg = 2;
    if ((sM <= R(0)) && (R(0) < sRs)) reg = 3;
    if ((sRs <= R(0)) && (R(0) <= sR)) reg = 4;
    if (sR < R(0)) reg = 5;
    V8<R> F;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        R v;
        switch (reg) {
            case 1: v = s.fm.c[i] + (QLs.c[i] - QL.c[i]) * sL; break;
            case 2: v = s.fm.c[i] + (QLss.c[i] - QLs.c[i]) * sLs; break;
            case 3: v = s.fp.c[i] + (QRss.c[i] - QRs.c[i]) * sRs; break;
            case 4: v = s.fp.c[i] + (QRs.c[i] - QR.c[i]) * sR; break;
            case 5: v = s.fp.c[i]; break;
            default: v = s.fm.c[i]; break;
        }
        F.c[i] = v;
    }
    return F;
}

// ---- Osher-Solomon flux with the Dumbser-Toro path integral (solvers.py:236-284)
namespace dots {
// 3-point Gauss-Legendre on [0, 1] as the reference builds them
// (generic.py:246-248: numpy leggauss(3), x/2 + 1/2, w/2), bit for bit
__device__ constexpr double ROOT[3] = {0x1.cda042f0236e0p-4, 0x1.0000000000000p-1, 0x1.c64bf7a1fb924p-1};
__device__ constexpr double WEIGHT[3] = {0x1.1c71c71c71c73p-2, 0x1.c71c71c71c71cp-2, 0x1.1c71c71c71c73p-2};

__device__ inline double sign1(double x) {  // np.sign with 0 -> 1 (constructor.py:198-199)
    return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : (x == 0.0 ? 1.0 : x));
}

// wave speeds of a state read as (rho, v, p, B): sound, Alfven along Bn,
// fast, slow (constructor.py:189-194, solvers.py:252-257)
struct Speeds {
    double cs, cax, cf, csl;
};
__device__ inline Speeds speeds(const double g[8], double Bn, double gamma) {
    const double rho = g[0], p = g[4];
    Speeds w;
    w.cs = ah_sqrt(gamma * gdiv(p, rho));
    const double ca = gdiv(ah_sqrt((g[5] * g[5] + g[6] * g[6]) + g[7] * g[7]), ah_sqrt(rho));
    w.cax = gdiv(Bn, ah_sqrt(rho));
    const double cs2 = w.cs * w.cs, ca2 = w.cax * w.cax, s = cs2 + ca * ca;
    const double disc = ah_sqrt(s * s - 4.0 * cs2 * ca2);
    w.cf = ah_sqrt(0.5 * (s + disc));
    w.csl = ah_sqrt(0.5 * (s - disc));
    return w;
}

// Right eigenvectors of adiabatic MHD at g (constructor.py:180-290), column
// major: Rc[k][i] = R[i][k].
__device__ inline void right_eigenvectors(int axis, const double g[8], double gamma, double Rc[8][8]) {
    const int a = axis % 3, o = (axis + 1) % 3, c = (axis + 2) % 3;
    const double rho = g[0];
    const double vx = g[1 + a], vy = g[1 + o], vz = g[1 + c];
    const double Bx = g[5 + a], By = g[5 + o], Bz = g[5 + c];
    const Speeds w = speeds(g, Bx, gamma);
    const double cs = w.cs, cf = w.cf, csl = w.csl;
    const double S = sign1(Bx);
    const double cs2 = cs * cs, cf2 = cf * cf, csl2 = csl * csl;
    const double af = ah_sqrt(gdiv(cs2 - csl2, cf2 - csl2));
    const double als = ah_sqrt(gdiv(cf2 - cs2, cf2 - csl2));
    const double bt2 = By * By + Bz * Bz;
    const double bperp = ah_sqrt(gdiv(bt2, rho));
    const double b2 = gdiv(By, ah_sqrt(bt2)), b3 = gdiv(Bz, ah_sqrt(bt2));
    const double vn = ah_sqrt((g[1] * g[1] + g[2] * g[2]) + g[3] * g[3]);
    const double v2 = vn * vn;
    const double tr = vy * b2 + vz * b3;
    const double base_s = (0.5 * als * rho * v2 - cs * af * rho * bperp) + (als * rho * cs2) / (gamma - 1.0);
    const double ts4 = als * csl * rho * vx, ts5 = af * cf * rho * S * tr;
    const double base_f = (0.5 * af * rho * v2 + cs * als * rho * bperp) + (af * rho * cs2) / (gamma - 1.0);
    const double tf4 = af * cf * rho * vx, tf5 = als * csl * rho * S * tr;
    const double r15 = pow(rho, 1.5), sr = ah_sqrt(rho);
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int i = 0; i < 8; ++i) Rc[k][i] = 0.0;
    // fast -/+ (columns 0, 7)
    Rc[0][0] = rho * af;
    Rc[0][a + 1] = rho * af * (vx - cf);
    Rc[0][o + 1] = rho * (af * vy + als * csl * b2 * S);
    Rc[0][c + 1] = rho * (af * vz + als * csl * b3 * S);
    Rc[0][4] = (base_f - tf4) + tf5;
    Rc[0][o + 5] = als * cs * b2 * sr;
    Rc[0][c + 5] = als * cs * b3 * sr;
    Rc[7][0] = rho * af;
    Rc[7][a + 1] = rho * af * (vx + cf);
    Rc[7][o + 1] = rho * (af * vy - als * csl * b2 * S);
    Rc[7][c + 1] = rho * (af * vz - als * csl * b3 * S);
    Rc[7][4] = (base_f + tf4) - tf5;
    Rc[7][o + 5] = als * cs * b2 * sr;
    Rc[7][c + 5] = als * cs * b3 * sr;
    // slow -/+ (columns 2, 5)
    Rc[2][0] = rho * als;
    Rc[2][a + 1] = rho * als * (vx - csl);
    Rc[2][o + 1] = rho * (als * vy - af * cf * b2 * S);
    Rc[2][c + 1] = rho * (als * vz - af * cf * b3 * S);
    Rc[2][4] = (base_s - ts4) - ts5;
    Rc[2][o + 5] = -af * cs * b2 * sr;
    Rc[2][c + 5] = -af * cs * b3 * sr;
    Rc[5][0] = rho * als;
    Rc[5][a + 1] = rho * als * (vx + csl);
    Rc[5][o + 1] = rho * (als * vy + af * cf * b2 * S);
    Rc[5][c + 1] = rho * (als * vz + af * cf * b3 * S);
    Rc[5][4] = (base_s + ts4) + ts5;
    Rc[5][o + 5] = -af * cs * b2 * sr;
    Rc[5][c + 5] = -af * cs * b3 * sr;
    // Alfven -/+ (columns 1, 6)
    Rc[1][o + 1] = -b3 * r15;
    Rc[1][c + 1] = b2 * r15;
    Rc[1][4] = (b2 * vz - b3 * vy) * r15;
    Rc[1][o + 5] = -rho * b3;
    Rc[1][c + 5] = rho * b2;
    Rc[6][o + 1] = b3 * r15;
    Rc[6][c + 1] = -b2 * r15;
    Rc[6][4] = (b3 * vy - b2 * vz) * r15;
    Rc[6][o + 5] = -rho * b3;
    Rc[6][c + 5] = rho * b2;
    // entropy and divergence waves (columns 3, 4)
    Rc[3][0] = 1.0;
    Rc[3][a + 1] = vx;
    Rc[3][o + 1] = vy;
    Rc[3][c + 1] = vz;
    Rc[3][4] = 0.5 * v2;
    Rc[4][4] = Bx;
    Rc[4][a + 5] = 1.0;
}

// One-sided (Hestenes) Jacobi SVD of the 8x8 matrix with columns A[k]:
// on return the columns of A are orthogonal (A_final = A V) and V holds the
// accumulated rotations, so A = U diag(sigma) V^T with sigma_k = |A_final[k]|.
__device__ inline void jacobi_svd8(double A[8][8], double V[8][8]) {
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int i = 0; i < 8; ++i) V[k][i] = i == k ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 30; ++sweep) {
        bool rotated = false;
#pragma unroll
        for (int p = 0; p < 7; ++p)
#pragma unroll
            for (int q = p + 1; q < 8; ++q) {
                double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    al += A[p][i] * A[p][i];
                    be += A[q][i] * A[q][i];
                    ga += A[p][i] * A[q][i];
                }
                if (!(ah_fabs(ga) > 1e-15 * ah_sqrt(al * be))) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (ah_fabs(zeta) + ah_sqrt(1.0 + zeta * zeta));
                const double cs = 1.0 / ah_sqrt(1.0 + t * t), sn = cs * t;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const double x = A[p][i], y = A[q][i];
                    A[p][i] = cs * x - sn * y;
                    A[q][i] = sn * x + cs * y;
                    const double u = V[p][i], v = V[q][i];
                    V[p][i] = cs * u - sn * v;
                    V[q][i] = sn * u + cs * v;
                }
            }
        if (!rotated) break;
    }
}
}  // namespace dots

// Osher-Solomon flux (solvers.py:236-284).  The path psi(s) = q- + s (q+ - q-)
// through the conservative states is handed to the eigenvector routine and the
// wave speeds as if primitive (244-252); the eigenvalues use component 1 and
// B_x whatever the axis (251-266), the eigenvectors the solver axis.
// |A| = R |Lambda| pinv(R) (271) with pinv from the SVD, singular values at
// or below 1e-15 sigma_max dropped as numpy's pinv does; the flux is
// (F+ + F-)/2 - (J q+ - J q-)/2 with J the weighted sum over the nodes.
// All 8 components enter, as in the reference (dead ones are zero).
template <int MASK>
__device__ inline V8<double> flux_dots(const Iface<double>& s, int sax, double gamma) {
    double xp[8], xm[8];
    bool finite = true;
#pragma unroll
    for (int i = 0; i < 8; ++i) xp[i] = xm[i] = 0.0;
    for (int node = 0; node < 3; ++node) {
        double g[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) g[i] = s.qm.c[i] + dots::ROOT[node] * (s.qp.c[i] - s.qm.c[i]);
        double A[8][8], V[8][8];
        dots::right_eigenvectors(sax, g, gamma, A);
        // numpy's pinv raises (its SVD does not converge) on a matrix with a
        // non-finite entry: the reference's run ends inside this step
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int c = 0; c < 8; ++c) finite = finite && ah_isfinite(A[r][c]);
        const dots::Speeds w = dots::speeds(g, g[5], gamma);
        const double vx = g[1];
        const double lam[8] = {ah_fabs(vx - w.cf),  ah_fabs(vx - w.cax), ah_fabs(vx - w.csl), ah_fabs(vx),
                               ah_fabs(vx),         ah_fabs(vx + w.csl), ah_fabs(vx + w.cax), ah_fabs(vx + w.cf)};
        dots::jacobi_svd8(A, V);
        double sig2[8], smax = 0.0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            double n = 0.0;
#pragma unroll
            for (int i = 0; i < 8; ++i) n += A[k][i] * A[k][i];
            sig2[k] = n;
            smax = nmax(smax, ah_sqrt(n));
        }
        // pinv(R) q = V diag(1/sigma^2) A_final^T q, then |A| q = R |Lambda| (that)
        double yp[8], ym[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const bool keep = ah_sqrt(sig2[k]) > 1e-15 * smax;
            double dp = 0.0, dm = 0.0;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                dp += A[k][i] * s.qp.c[i];
                dm += A[k][i] * s.qm.c[i];
            }
            yp[k] = keep ? dp / sig2[k] : 0.0;
            ym[k] = keep ? dm / sig2[k] : 0.0;
        }
        double zp[8], zm[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            double ap = 0.0, am = 0.0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                ap += V[k][j] * yp[k];
                am += V[k][j] * ym[k];
            }
            zp[j] = lam[j] * ap;
            zm[j] = lam[j] * am;
        }
        dots::right_eigenvectors(sax, g, gamma, A);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            double ap = 0.0, am = 0.0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                ap += A[k][i] * zp[k];
                am += A[k][i] * zm[k];
            }
            xp[i] += dots::WEIGHT[node] * ap;
            xm[i] += dots::WEIGHT[node] * am;
        }
    }
    V8<double> F = zero8<double, MASK>();
#pragma unroll
    for (int i = 0; i < 8; ++i)
        if (VS<MASK>::live(i)) F.c[i] = finite ? 0.5 * (s.fp.c[i] + s.fm.c[i]) - 0.5 * (xp[i] - xm[i]) : NAN;
    return F;
}

// ---- entropy-stable flux (solvers.py:288-391), 1D as in the reference --------
namespace esf {
__device__ inline double am(double l, double r) { return 0.5 * (l - r); }  // "arith_mean" (306-307)
__device__ inline double glog(double x) { return x > 0.0 ? log(x) : 0.0; }   // fv.log (fv.py:25-26)
__device__ inline double lon(double l, double r) { return gdiv(l - r, glog(l) - glog(r)); }
}  // namespace esf

// Entropy-conserving flux (Winters & Gassner) plus the hybrid Roe/LLF
// dissipation R |Lambda| R^T (entropy jump) of Derigs et al., as the
// reference writes them: the "arithmetic mean" is the half-difference
// .5 (L - R), velocity and field components are taken at indices 0, 1, 2
// whatever the axis, the eigenvectors (constructor.py:294-415) at the
// averaged state with the axis.  Every matrix product is evaluated term by
// term, so a non-finite entry spreads along its row exactly as in the
// reference's matmuls (most of its outputs are non-finite).
template <int MASK>
__device__ inline V8<double> flux_es(const Iface<double>& s, int sax, double gamma) {
    using esf::am;
    using esf::lon;
    const int a = sax % 3, o = (sax + 1) % 3, c = (sax + 2) % 3;
    const V8<double>& L = s.wm;
    const V8<double>& Rr = s.wp;
    const double rhoL = L.c[0], PL = L.c[4], rhoR = Rr.c[0], PR = Rr.c[4];
    const double z1L = ah_sqrt(gdiv(rhoL, PL)), z1R = ah_sqrt(gdiv(rhoR, PR));
    const double z5L = ah_sqrt(rhoL * PL), z5R = ah_sqrt(rhoR * PR);
    const double rho_h = am(z1L, z1R) * lon(z5L, z5R);
    const double P1 = gdiv(am(z5L, z5R), am(z1L, z1R));
    const double P2 = ((gamma + 1.0) / (2.0 * gamma)) * (gdiv(lon(z5L, z5R), lon(z1L, z1R))) +
                      ((gamma - 1.0) / (2.0 * gamma)) * (gdiv(am(z5L, z5R), am(z1L, z1R)));
    double hv1[3], hv2[3], Bh[3], Bd[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double vl = L.c[1 + k], vr = Rr.c[1 + k];
        hv1[k] = gdiv(am(vl * z1L, vr * z1R), am(z1L, z1R));
        hv2[k] = gdiv(am(vl * (z1L * z1L), vr * (z1R * z1R)), am(z1L * z1L, z1R * z1R));
        Bh[k] = am(L.c[5 + k], Rr.c[5 + k]);
        Bd[k] = am(L.c[5 + k] * L.c[5 + k], Rr.c[5 + k] * Rr.c[5 + k]);
    }
    const double B1B2 = am(L.c[5] * L.c[6], Rr.c[5] * Rr.c[6]), B1B3 = am(L.c[5] * L.c[7], Rr.c[5] * Rr.c[7]);
    const double u1 = hv1[0], v1 = hv1[1], w1 = hv1[2], u2 = hv2[0], v2 = hv2[1], w2 = hv2[2];
    double ec[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) ec[i] = 0.0;
    ec[0] = rho_h * u1;
    ec[a + 1] = P1 + rho_h * (u1 * u1) + 0.5 * ((Bd[0] + Bd[1]) + Bd[2]) - Bd[0];
    ec[o + 1] = rho_h * u1 * v1 - B1B2;
    ec[c + 1] = rho_h * u1 * w1 - B1B3;
    ec[4] = (gamma / (gamma - 1.0)) * u1 * P2 + 0.5 * rho_h * u1 * ((u1 * u1 + v1 * v1) + w1 * w1) +
            u2 * (Bh[1] * Bh[1] + Bh[2] * Bh[2]) - Bh[0] * (v2 * Bh[1] + w2 * Bh[2]);
    ec[o + 5] = u2 * Bh[1] - v2 * Bh[0];
    ec[c + 5] = u2 * Bh[2] - w2 * Bh[0];
    // eigenvectors at the averaged state, then R = R0 @ sqrt(D) (constructor.py:404-414)
    const double hat[8] = {rho_h, u1, v1, w1, P1, Bh[0], Bh[1], Bh[2]};
    double R0[8][8];  // column major
    dots::right_eigenvectors(sax, hat, gamma, R0);
    double Dd[8];
    Dd[0] = 1.0 / (2.0 * gamma * rho_h);
    Dd[a + 1] = gdiv(P1, 2.0 * (rho_h * rho_h));
    Dd[o + 1] = 1.0 / (2.0 * gamma * rho_h);
    Dd[c + 1] = (rho_h * (gamma - 1.0)) / gamma;
    Dd[4] = gdiv(P1, rho_h);
    Dd[a + 5] = 1.0 / (2.0 * gamma * rho_h);
    Dd[o + 5] = gdiv(P1, 2.0 * (rho_h * rho_h));
    Dd[c + 5] = 1.0 / (2.0 * gamma * rho_h);
    double Rm[8][8];  // row major: Rm[i][j] = sum_k R0[i][k] sqrt(D)[k][j]
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 8; ++k) acc += R0[k][i] * (k == j ? ah_sqrt(Dd[j]) : 0.0);
            Rm[i][j] = acc;
        }
    // wave speeds at the averaged state and the hybrid eigenvalue matrix
    double Bhat[8] = {0.0, 0.0, 0.0, 0.0, 0.0, Bh[0], Bh[1], Bh[2]};
    Bhat[0] = rho_h;
    Bhat[4] = P1;
    const dots::Speeds w = dots::speeds(Bhat, Bh[0], gamma);
    const double wv[8] = {u1 + w.cf, u1 + w.cax, u1 + w.csl, u1, u1, u1 - w.csl, u1 - w.cax, u1 - w.cf};
    double mx = ah_fabs(wv[0]);
#pragma unroll
    for (int k = 1; k < 8; ++k)
        if (k != 4) {
            const double x = ah_fabs(wv[k]);
            mx = (x > mx || x != x) ? x : mx;  // np.maximum.reduce: NaN propagates
        }
    const double eps = ah_sqrt(ah_fabs(gdiv(PR - PL, PR + PL)));
    const double offd = (1.0 - eps) * 0.0 + eps * 0.0;
    double lamd[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) lamd[k] = (1.0 - eps) * ah_fabs(wv[k]) + eps * mx;
    // entropy-variable jump (solvers.py:371-377), negated
    const double nL = ah_sqrt((L.c[1] * L.c[1] + L.c[2] * L.c[2]) + L.c[3] * L.c[3]);
    const double nR = ah_sqrt((Rr.c[1] * Rr.c[1] + Rr.c[2] * Rr.c[2]) + Rr.c[3] * Rr.c[3]);
    double e[8];
    e[0] = ((gamma - log(PL * pow(rhoL, -gamma))) / (gamma - 1.0) - gdiv(0.5 * rhoL * (nL * nL), PL)) -
           ((gamma - log(PR * pow(rhoR, -gamma))) / (gamma - 1.0) - gdiv(0.5 * rhoR * (nR * nR), PR));
    e[4] = gdiv(rhoR, PR) - gdiv(rhoL, PL);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        e[1 + k] = gdiv(L.c[1 + k] * rhoL, PL) - gdiv(Rr.c[1 + k] * rhoR, PR);
        e[5 + k] = gdiv(L.c[5 + k] * rhoL, PL) - gdiv(Rr.c[5 + k] * rhoR, PR);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) e[k] = -e[k];
    // dissipation = (R lam) R^T e, streamed by rows of R lam
    V8<double> F = zero8<double, MASK>();
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        double M[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            double acc = 0.0;
#pragma unroll
            for (int m = 0; m < 8; ++m) acc += Rm[i][m] * (m == k ? lamd[k] : offd);
            M[k] = acc;
        }
        double d = 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            double aij = 0.0;
#pragma unroll
            for (int k = 0; k < 8; ++k) aij += M[k] * Rm[j][k];
            d += aij * e[j];
        }
        if (VS<MASK>::live(i)) F.c[i] = ec[i] + 0.5 * d;
    }
    return F;
}

}  // namespace ah

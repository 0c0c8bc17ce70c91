// Kernels of the astrea finite-volume update on gfx950 (wave64, 256-thread
// workgroups).  See ah_device.h for the per-cell arithmetic.
//
// Index conventions (reference Appendix "sweep frame", schemes/*.py):
//   s = index along the sweep axis AX, t = index along the other axis.
//   Interface k in [0, n_s] separates cells k-1 and k.
//   w_minus[k] = wR[bc(k-1)], w_plus[k] = wL[bc(k)] (schemes/plm.py:81).
//   Every array the reference pads (fv.add_boundary, functions/fv.py:57-61) is
//   evaluated at e(p) = clamp(p, 0, n-1) for edge boundaries; for periodic
//   boundaries values are computed in place from wrapped loads (translation
//   invariance), so e(p) = p inside a tile window.
#pragma once
#include "ah_device.h"

#include <type_traits>

namespace ah {

struct ProbDesc {
    int n0, n1, bc, nprob_pad;
    long long off;
    double dx, gamma, cfl, t_end, chkpt;
    long long max_steps;
};

struct ProbState {
    double t, dt;
    long long steps;
    int idx, active, failed, parity;
    unsigned long long eig[2];  // bit pattern of max spectral radius per axis (stage 0)
    int fail_flag;
    int ckpt;  // 1 when this step's dt was clamped to the next checkpoint (astrea.py:74-78)
    // PPM authors c/ph: the reference's grid-wide interface-limiter tests of
    // this step, two bits per stage (bit 2k: some face wF is a local extremum,
    // bit 2k+1: some left-face value wF_L is), cleared when a step ends
    unsigned long long aflags;
};

struct Tile {
    int prob, s0, t0, pad;
};

// This step's dt from the stage-0 spectral radii (astrea.py:70-78): the CFL
// value, clamped to the next checkpoint (*hit = 1 then).  A pure function of
// the step's time state, so every kernel of the step evaluates the same dt.
__device__ inline double step_dt(const ProbDesc& d, const ProbState& s, int dims, int* hit) {
#pragma clang fp contract(off)
    const double e0 = __longlong_as_double((long long)s.eig[0]);
    const double e1 = dims == 2 ? __longlong_as_double((long long)s.eig[1]) : e0;
    const double a0 = d.dx / e0, a1 = d.dx / e1;
    double dt = d.cfl * (dims == 2 ? (a1 < a0 ? a1 : a0) : a0);
    *hit = 0;
    if (s.t + dt >= d.chkpt * s.idx) {
        dt = d.chkpt * s.idx - s.t;
        *hit = 1;
    }
    return dt;
}

struct SweepArgs {
    const double* X;      // stage input state, packed SoA (plane stride `plane`)
    const double* xl;     // non-null: the stage input is X + dt xl, the first SSPRK(3,3)
                          // register K1 = U + dt L0 evaluated where it is read
    double* out;          // 1D: L;  2D flux: D (first sweep) or L (second)
    const double* Dprev;  // 2D second sweep: D of the first sweep
    void* wL;             // 2D interface states, sweep-frame packed SoA, type R
    void* wR;
    void* bn;             // 2D HLLD: centred normal field per cell (type R)
    void* wF;             // 2D MHD: the sweep's face values feeding the corner EMF (type R)
    double* flux_out;     // optional: interface fluxes (parity entry point)
    long long plane;
    long long flux_plane;
    const ProbDesc* pd;
    ProbState* ps;
    const Tile* tiles;
    int ntiles;             // tile count when ntiles_dev is null
    const int* ntiles_dev;  // device-resident tile count (compactable lists)
    int stage0;   // accumulate the eigmax for dt (stage 0 of a step)
    int limiter;  // PLM slope limiter
    int flattener;
    int parity_override;  // -1: use per-problem parity
    int last;             // 2D: this flux launch finishes L
    int compat;           // 0: reference solver axis (enumeration index), 1: physical axis
    int av;               // 1D PPM: add the artificial-viscosity flux alpha nu (q_k - q_{k-1})
    int stage_idx;        // stage of the step (PPM authors c/ph read that stage's limiter tests)
    int flags_only;       // PPM authors c/ph: only record this stage's interface-limiter tests
    // 2D second sweep of the middle SSPRK(3,3) stage: write K2 = (3U + K1 + dt L) / 4
    // (evolvers.py:172) in place of L, in the combine kernel's operation order
    double* fuse_k2;      // null: write L
    const double* fuse_u;
    const double* fuse_k1;
};

// Stage input at packed index j: X[j], or K1 = U + dt L0 (evolvers.py:169)
// in the stage combination's operation order when the register is not stored
__device__ inline double stage_x(const SweepArgs& a, long long j, double xdt) {
#pragma clang fp contract(off)
    if (a.xl) return a.X[j] + xdt * a.xl[j];
    return a.X[j];
}

__device__ inline int clampi(int p, int lo, int hi) { return p < lo ? lo : (p > hi ? hi : p); }
__device__ inline int bcmap(int p, int n, int bc) {
    if (bc == WRAP) {
        int r = p % n;
        return r < 0 ? r + n : r;
    }
    return clampi(p, 0, n - 1);
}

__device__ inline void atomic_max_pos(unsigned long long* a, double v) {
    // non-negative doubles order like their bit patterns; NaN bit patterns are
    // larger than +inf so a NaN wins and flags the failure.
    atomicMax(a, (unsigned long long)__double_as_longlong(v));
}

template <class R, int MASK>
__device__ inline V8<R> load8(const R* base, long long stride, int idx) {
    using V = VS<MASK>;
    V8<R> w = zero8<R, MASK>();
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (V::live(c)) w.c[c] = base[V::slot(c) * stride + idx];
    return w;
}
template <class R, int MASK>
__device__ inline void store8(R* base, long long stride, int idx, const V8<R>& w) {
    using V = VS<MASK>;
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (V::live(c)) base[V::slot(c) * stride + idx] = w.c[c];
}
template <class D, class S, int MASK>
__device__ inline V8<D> cast8(const V8<S>& a) {
    V8<D> b;
#pragma unroll
    for (int c = 0; c < 8; ++c) b.c[c] = VS<MASK>::live(c) ? D(a.c[c]) : D(0);
    return b;
}

// 4th-order cell-average conversion q -> w (functions/fv.py:126-143), FP64:
// w = C(q - K24 lap0 q - K24 lap1 q) + (K24 lap0 C(q) + K24 lap1 C(q)),
// C = pointwise cons->prim.  lap_a x = (x[+1] - x[0]) - (x[0] - x[-1]).
template <int MASK, int DIM, bool SPLIT = false>
__device__ inline V8<double> ho_cell_prim(const V8<double>& q, const V8<double>& qs0m, const V8<double>& qs0p,
                                          const V8<double>& qt0m, const V8<double>& qt0p, const V8<double>& cq,
                                          const V8<double>& cs0m, const V8<double>& cs0p, const V8<double>& ct0m,
                                          const V8<double>& ct0p, double gamma) {
    using V = VS<MASK>;
    V8<double> qq = q, w = zero8<double, MASK>();
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        if (!V::live(c)) continue;
        // axis order of the reference loop (ax = 0 then 1) in the array the
        // conversion receives, i.e. the sweep frame: first s then t.
        qq.c[c] = q.c[c] - K24 * ((qs0p.c[c] - q.c[c]) - (q.c[c] - qs0m.c[c]));
        double acc = K24 * ((cs0p.c[c] - cq.c[c]) - (cq.c[c] - cs0m.c[c]));
        if constexpr (DIM == 2) {
            qq.c[c] = qq.c[c] - K24 * ((qt0p.c[c] - q.c[c]) - (q.c[c] - qt0m.c[c]));
            acc = acc + K24 * ((ct0p.c[c] - cq.c[c]) - (cq.c[c] - ct0m.c[c]));
        }
        w.c[c] = acc;
    }
    V8<double> r = SPLIT ? cons_to_prim_split<MASK>(qq, gamma) : cons_to_prim<double, MASK>(qq, gamma);
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (V::live(c)) r.c[c] = r.c[c] + w.c[c];
    return r;
}

// The flattener's chi~ (ppm.py:111-124) at position mm, from the pressure at
// e(mm)-2..e(mm)+2 and the normal velocity at e(mm)-1, e(mm)+1.
template <class R, int AX, class Acc>
__device__ inline R flat_chi_at(const Acc& W, int flattener, int mm) {
    const int e2 = W.e(mm);
    const auto p2 = W(e2 - 2), p1 = W(e2 - 1), q1 = W(e2 + 1), q2 = W(e2 + 2);
    return flat_chi<R>(flattener, p2.c[4], p1.c[4], q1.c[4], q2.c[4], p1.c[AX + 1], q1.c[AX + 1]);
}

// One cell's reconstruction from a window accessor W(m) that returns the
// primitive state at sweep position m (already boundary mapped).  `ce` is the
// evaluation position of the cell; chi_at(mm) returns flat_chi_at(W, mm)
// (computed, or read from a table of the row).
template <class R, int MASK, int RECON, int AX, class Acc, class ChiF>
__device__ inline void recon_cell(const Acc& W, int ce, int limiter, int flattener, V8<R>& wL, V8<R>& wR,
                                  const ChiF& chi_at) {
    using V = VS<MASK>;
    wL = zero8<R, MASK>();
    wR = zero8<R, MASK>();
    if constexpr (RECON == PCM) {
        wL = wR = W(ce);
    } else if constexpr (RECON == PLM) {
        const V8<R> a = W(ce - 1), z = W(ce), b = W(ce + 1);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (!V::live(c)) continue;
            const R g = R(0.5) * slope_limit(limiter, z.c[c] - a.c[c], b.c[c] - z.c[c]);
            wL.c[c] = z.c[c] - g;
            wR.c[c] = z.c[c] + g;
        }
    } else if constexpr (RECON == PPM) {
        // The flattener's eta at the two faces depends on pressure and the
        // normal velocity only; it is computed once, then every component runs
        // its own chain (interpolant, curvatures, third differences, limiter),
        // which keeps one component's window in registers at a time.  Every
        // padded intermediate at position m is evaluated at W.e(m).
        R etaL = R(1), etaR = R(1);
        if (flattener != FLAT_OFF) {
            // eta(me) = min(chi(me), chi(me -/+ 1)) by sign of the pressure jump (ppm.py:127-131)
            auto eta_at = [&](int m) {
                const int me = W.e(m);
                const R chi0 = chi_at(me);
                const R sg = nsign(W(me + 1).c[4] - W(me - 1).c[4]);
                R eta = chi0;
                if (sg < R(0)) eta = nmin(chi0, chi_at(me + 1));
                if (sg > R(0)) eta = nmin(chi0, chi_at(me - 1));
                return eta;
            };
            etaL = eta_at(ce - 1);
            etaR = eta_at(ce);
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (!V::live(c)) continue;
            auto w = [&](int m) { return W(m).c[c]; };
            // face interpolant at me (ppm.py:38), blended toward w(me) by eta (ppm.py:65-67)
            auto wF_at = [&](int m, R eta) {
                const int me = W.e(m);
                const R b = w(me);
                R f = R(K7_12) * (b + w(me + 1)) - R(K1_12) * (w(me - 1) + w(me + 2));
                if (flattener != FLAT_OFF) f = f * eta + b * (R(1) - eta);
                return f;
            };
            auto d2_at = [&](int m) {
                const int me = W.e(m);
                return w(me - 1) - R(2) * w(me) + w(me + 1);
            };
            // d3w[m] = d2w_C[e(m+1)] - d2w_C[m] (limiters.py:96), padded by two (126-127)
            auto d3_at = [&](int m) {
                const int me = W.e(m);
                return d2_at(me + 1) - d2_at(me);
            };
            const R fL = wF_at(ce - 1, etaL), fR = wF_at(ce, etaR);
            // curvatures d2w_C padded by one (limiters.py:93,103): at e(ce-1), ce, e(ce+1)
            const R d2m = d2_at(ce - 1), d2c = d2_at(ce), d2p = d2_at(ce + 1);
            const R d3m2 = d3_at(ce - 2), d3m1 = d3_at(ce - 1), d30 = d3_at(ce), d3p2 = d3_at(ce + 2);
            mc_limit<R>(w(ce), w(ce - 1), w(ce + 1), w(ce - 2), w(ce + 2), fL, fR, d2m, d2c, d2p, d3m2, d3m1, d30,
                        d3p2, wL.c[c], wR.c[c]);
        }
    } else {
        constexpr int H = RECON == WENO3 ? 1 : (RECON == WENO5 ? 2 : 3);
        V8<R> w[7];
#pragma unroll
        for (int k = 0; k < 2 * H + 1; ++k) w[k] = W(ce - H + k);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (!V::live(c)) continue;
            if constexpr (RECON == WENO3)
                weno3<R>(w[0].c[c], w[1].c[c], w[2].c[c], wL.c[c], wR.c[c]);
            else if constexpr (RECON == WENO5)
                weno5<R>(w[0].c[c], w[1].c[c], w[2].c[c], w[3].c[c], w[4].c[c], wL.c[c], wR.c[c]);
            else
                weno7<R>(w[0].c[c], w[1].c[c], w[2].c[c], w[3].c[c], w[4].c[c], w[5].c[c], w[6].c[c], wL.c[c],
                         wR.c[c]);
        }
    }
}

template <class R, int MASK, int RECON, int AX, class Acc>
__device__ inline void recon_cell(const Acc& W, int ce, int limiter, int flattener, V8<R>& wL, V8<R>& wR) {
    recon_cell<R, MASK, RECON, AX>(W, ce, limiter, flattener, wL, wR,
                                   [&](int mm) { return flat_chi_at<R, AX>(W, flattener, mm); });
}

// ---- PPM authors 'c' and 'ph' (ppm.py:43-62, 69-71) ----------------------------
// Colella et al. 2011 and Peterson & Hammett 2008: not reachable from the
// reference's options (evolvers.py:17 passes author 'mc'), built as the
// ppm_author option for 1D hydrodynamics.
constexpr int AUTH_MC = 0, AUTH_C = 1, AUTH_PH = 2;

// limiters.py:53-79 for one face value wf between cells w0 and wp1; `any`
// is the reference's grid-wide test (line 58): some face of the sweep array
// is a local extremum, (wf - w0)(wp1 - wf) < 0
template <class R>
__device__ inline R iface_limit(bool any, R wf, R wm1, R w0, R wp1, R wp2) {
    if (!any) return wf;
    constexpr double C = 5.0 / 4.0;
    const R dl = wm1 - R(2) * w0 + wp1;
    const R dc = R(3) * (w0 - R(2) * wf + wp1);
    const R dr = w0 - R(2) * wp1 + wp2;
    const R sl = nsign(dl), sc = nsign(dc), sr = nsign(dr);
    R d2 = R(0);
    if (sl == sr && sc == sr && sc == sl)
        d2 = sc * nmin(ah_fabs(dc), nmin(ah_fabs(R(C) * dl), ah_fabs(R(C) * dr)));
    return R(0.5) * (w0 + wp1) - d2 / R(6);
}

template <class R>
__device__ inline bool iface_extremum(R wf, R w0, R wp1) { return (wf - w0) * (wp1 - wf) < R(0); }

// One cell's extrapolants for author c / ph (limiters.py:144-201), component
// by component.  The reference's grid-wide branch test of line 163 is true on
// every array this build runs these authors on: its 8-component arrays carry
// the components a 1D hydrodynamic state keeps at zero, whose
// (wS - wF_L)(wF_R - wS) = 0 <= 0 marks every cell as an extremum.
// anyR / anyL: the interface-limiter tests of this stage for the face values
// wF (faces i+1/2) and, for 'ph', the left-face values wF_L.
template <class R, int MASK, int AUTH, class Acc>
__device__ inline void author_cell(const Acc& W, int ce, bool anyR, bool anyL, V8<R>& wL, V8<R>& wR) {
    using V = VS<MASK>;
    constexpr double C = 5.0 / 4.0;
    wL = zero8<R, MASK>();
    wR = zero8<R, MASK>();
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        if (!V::live(c)) continue;
        auto w = [&](int m) { return W(m).c[c]; };
        // face m+1/2 (ppm.py:38) at face index e(m), limited (limiters.py:53-79)
        auto face = [&](int m) {
            const int me = W.e(m);
            const R f = R(K7_12) * (w(me) + w(me + 1)) - R(K1_12) * (w(me - 1) + w(me + 2));
            return iface_limit<R>(anyR, f, w(me - 1), w(me), w(me + 1), w(me + 2));
        };
        const R s0 = w(ce), sm = w(ce - 1), sp = w(ce + 1), sm2 = w(ce - 2), sp2 = w(ce + 2);
        R fL, fR;
        if constexpr (AUTH == AUTH_PH) {
            // ppm.py:51-57: the cell's own left-face interpolant and its face i+1/2
            const R fl0 = R(K7_12) * (sm + s0) - R(K1_12) * (sm2 + sp);
            fL = iface_limit<R>(anyL, fl0, sm2, sm, s0, sp);
            const R fr0 = R(K7_12) * (s0 + sp) - R(K1_12) * (sm + sp2);
            fR = iface_limit<R>(anyR, fr0, sm, s0, sp, sp2);
        } else {
            fL = face(ce - 1);
            fR = face(ce);
        }
        const R dm = s0 - fL, dp = fR - s0;
        const bool cell_ext = dm * dp <= R(0);
        bool ext_ext;
        if constexpr (AUTH == AUTH_PH) {
            ext_ext = (sm - s0) * (s0 - sp) <= R(0);
        } else {
            // limiters.py:155-161 (wF_pad2[:-4], wF_pad2[4:]: faces i-3/2 and i+5/2)
            const R dfl = fL - face(ce - 2), dfr = face(ce + 2) - fR;
            const R dsl = s0 - sm, dsr = sp - s0;
            const R df = nmin(ah_fabs(dfl), ah_fabs(dfr)), ds = nmin(ah_fabs(dsl), ah_fabs(dsr));
            ext_ext = ((df >= ds) && (dfl * dfr < R(0))) || ((ds >= df) && (dsl * dsr < R(0)));
        }
        const R d2 = R(6) * (fL - R(2) * s0 + fR);
        const R d2l = sm2 - R(2) * sm + s0;
        const R d2c = sm - R(2) * s0 + sp;
        const R d2r = s0 - R(2) * sp + sp2;
        const R g0 = nsign(d2), gc = nsign(d2c), gl = nsign(d2l), gr = nsign(d2r);
        const bool same = g0 == gc && g0 == gl && g0 == gr && gc == gl && gc == gr && gl == gr;
        const R lim = g0 * nmin(nmin(ah_fabs(d2), ah_fabs(R(C) * d2c)), nmin(ah_fabs(R(C) * d2l), ah_fabs(R(C) * d2r)));
        R d2lim = R(0);
        if (cell_ext && same) d2lim = lim;
        if constexpr (AUTH == AUTH_PH) {
            const R phi = gdiv(d2lim, d2);
            wL.c[c] = s0 + phi * (fL - s0);
            wR.c[c] = s0 + phi * (fR - s0);
        } else {
            if (ext_ext && same) d2lim = lim;
            const R phi = gdiv(d2lim, d2);
            const R dl = ah_fabs(dm) > R(2) * ah_fabs(dp) ? R(2) * dp : dm;
            const R dr = ah_fabs(dp) > R(2) * ah_fabs(dm) ? R(2) * dm : dp;
            wL.c[c] = s0 - phi * dl;
            wR.c[c] = s0 + phi * dr;
        }
    }
}

// this stage's interface-limiter tests of one cell's faces: bit 0 for its
// face i+1/2 (wF), bit 1 for its left-face value wF_L (author 'ph')
template <class R, int MASK, int AUTH, class Acc>
__device__ inline unsigned author_face_tests(const Acc& W, int ce) {
    using V = VS<MASK>;
    unsigned bits = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        if (!V::live(c)) continue;
        auto w = [&](int m) { return W(m).c[c]; };
        const R s0 = w(ce), sm = w(ce - 1), sp = w(ce + 1), sm2 = w(ce - 2), sp2 = w(ce + 2);
        const R fr0 = R(K7_12) * (s0 + sp) - R(K1_12) * (sm + sp2);
        if (iface_extremum<R>(fr0, s0, sp)) bits |= 1u;
        if constexpr (AUTH == AUTH_PH) {
            const R fl0 = R(K7_12) * (sm + s0) - R(K1_12) * (sm2 + sp);
            if (iface_extremum<R>(fl0, sm, s0)) bits |= 2u;
        }
    }
    return bits;
}

template <int MASK, bool SPLIT>
__device__ inline V8<double> c2p(const V8<double>& q, double gamma) {
    if constexpr (SPLIT) return cons_to_prim_split<MASK>(q, gamma);
    else return cons_to_prim<double, MASK>(q, gamma);
}
template <int MASK, bool SPLIT>
__device__ inline V8<double> p2c(const V8<double>& w, double gamma) {
    if constexpr (SPLIT) return prim_to_cons_split<MASK>(w, gamma);
    else return prim_to_cons<double, MASK>(w, gamma);
}
template <int MASK, int AX, bool SPLIT>
__device__ inline V8<double> pflux(const V8<double>& w, double gamma) {
    if constexpr (SPLIT) return phys_flux_split<MASK, AX>(w, gamma);
    else return phys_flux<double, MASK, AX>(w, gamma);
}
// spectral radius of a Jacobian state: FP32 in the split path (a coefficient)
template <int MASK, int AX, bool SPLIT>
__device__ inline double radius(const V8<double>& st, double gamma) {
    if constexpr (SPLIT) return double(spectral_radius<float, MASK, AX>(cast8<float, double, MASK>(st), float(gamma)));
    else return spectral_radius<double, MASK, AX>(st, gamma);
}
template <int MASK, bool SPLIT, bool ROE>
__device__ inline V8<double> jac_state(const V8<double>& wp, const V8<double>& wm) {
    if constexpr (SPLIT) {
        const V8<float> a = ROE ? roe_avg<float, MASK>(cast8<float, double, MASK>(wp), cast8<float, double, MASK>(wm))
                                : arith_avg<float, MASK>(cast8<float, double, MASK>(wp), cast8<float, double, MASK>(wm));
        return cast8<double, float, MASK>(a);
    } else {
        return ROE ? roe_avg<double, MASK>(wp, wm) : arith_avg<double, MASK>(wp, wm);
    }
}

template <class R, int MASK, int SOLVER, int AX>
__device__ inline V8<R> riemann(const Iface<R>& s, R lam, int sax, R Bn, R gamma, int mom, bool physical,
                               bool split = false) {
    if constexpr (SOLVER == LLF || SOLVER == LW)
        return flux_llf<R, MASK>(s, lam);
    else if constexpr (SOLVER == HLLC)
        return split ? flux_hllc_split<MASK>(s, sax, gamma, physical ? mom : 1, physical)
                     : flux_hllc<R, MASK>(s, sax, gamma, physical ? mom : 1, physical);
    else if constexpr (SOLVER == DOTS)
        return flux_dots<MASK>(s, sax, gamma);
    else if constexpr (SOLVER == ES)
        return flux_es<MASK>(s, sax, gamma);
    else
        return sax == 0 ? flux_hlld<R, MASK, 0>(s, Bn, gamma) : flux_hlld<R, MASK, 1>(s, Bn, gamma);
}

// Osher-Solomon: a non-finite flux marks the step failed (flux_dots returns
// NaN where the reference's pinv raises), which keeps the last good state
template <int SOLVER, int MASK, class R>
__device__ inline void dots_fail(const V8<R>& F, ProbState* ps) {
    if constexpr (SOLVER == DOTS) {
        bool nf = false;
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (VS<MASK>::live(c)) nf = nf || !ah_isfinite(F.c[c]);
        if (nf) atomicOr(&ps->fail_flag, 1);
    }
}

// Block-wide max of a non-negative double, then one atomic per block.
__device__ inline void block_max_atomic(double v, unsigned long long* dst, double* red) {
    // NaN-aware via bit patterns: reduce as unsigned long long
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(b, off, 64);
        b = o > b ? o : b;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    unsigned long long* r = reinterpret_cast<unsigned long long*>(red);
    if (lane == 0) r[wid] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = r[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = r[i] > m ? r[i] : m;
        if (m != 0ull) atomicMax(dst, m);
    }
}

// ============================================================================
// 1D: one fused kernel per stage: load -> cons->prim -> reconstruct/limit ->
// spectral radius -> Riemann -> flux difference -> L = -dF/dx.
// ============================================================================
// Window accessor of the 1D sweep: W(m) is the primitive state at e(m);
// CLAMP = false when the tile's stencil never leaves the domain or the
// boundary is periodic (e is then the identity).
template <class R, int MASK, int NW, bool CLAMP>
struct LineAcc {
    R (*ws)[NW];
    int bw, N;
    __device__ int e(int p) const {
        if constexpr (CLAMP) return clampi(p, 0, N - 1);
        else return p;
    }
    __device__ V8<R> operator()(int m) const {
        const int j = e(m) - bw;
        V8<R> v = zero8<R, MASK>();
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (VS<MASK>::live(c)) v.c[c] = ws[VS<MASK>::slot(c)][j];
        return v;
    }
};

template <class R, int MASK, int RECON, int SOLVER, int TS, bool SPLIT = false, int AUTH = AUTH_MC>
__global__ __launch_bounds__(256) void sweep1d_kernel(SweepArgs a) {
    static_assert(!SPLIT || sizeof(R) == 8, "split path keeps FP64 values");
    using V = VS<MASK>;
    constexpr int NV = V::NV;
    constexpr bool HO = (RECON == PPM || RECON >= WENO3);
    constexpr int HW_L = 5, HW_R = 6;                 // wS window [s0-5, s0+TS+6)
    constexpr int NQ = TS + HW_L + HW_R + 2;           // q window [s0-6, s0+TS+7)
    constexpr int NW = TS + HW_L + HW_R;
    constexpr int NC = TS + 3;                         // cells [s0-1, s0+TS+2)
    constexpr int NF = TS + 1;                         // interfaces [s0, s0+TS]
    constexpr int NR = TS + 2;                         // radii at [s0, s0+TS+1]
    __shared__ double qd[NV][NQ];
    __shared__ double cq[HO ? NV : 1][HO ? NQ : 1];
    __shared__ R ws[NV][NW];
    __shared__ R wl[RECON == PCM ? 1 : NV][RECON == PCM ? 1 : NC];
    __shared__ R wr[RECON == PCM ? 1 : NV][RECON == PCM ? 1 : NC];
    __shared__ R rad[NR];
    __shared__ R fl[NV][NF];
    __shared__ double red[4];
    __shared__ unsigned abits;  // PPM authors c/ph: the tile's interface-limiter tests
    // PPM flattener: chi~ at positions [s0-3, s0+TS+2]
    constexpr int NCHI = TS + 6;
    __shared__ R chi_s[RECON == PPM ? NCHI : 1];

    const int ntiles_ = wave_uniform(a.ntiles_dev ? *a.ntiles_dev : a.ntiles);
    for (int ti_ = blockIdx.x; ti_ < ntiles_; ti_ += gridDim.x) {
    const Tile tile = ld_const(a.tiles + ti_);
    const ProbDesc pd = ld_const(a.pd + tile.prob);
    ProbState* ps = a.ps + tile.prob;
    if (!ld_const(&ps->active)) continue;
    const int N = pd.n0, bc = pd.bc, s0 = tile.s0;
    const double gamma = pd.gamma;
    const R gR = R(gamma);
    const long long off = pd.off, P = a.plane;
    const int sax = 0;
    auto e = [&](int p) { return bc == EDGE ? clampi(p, 0, N - 1) : p; };
    const int bq = s0 - HW_L - 1, bw = s0 - HW_L, bcl = s0 - 1;
    int xhit_;
    const double xdt = a.xl ? step_dt(pd, *ps, 1, &xhit_) : 0.0;

    // A: load the stage input (conservative) with boundary mapping
    for (int i = threadIdx.x; i < NQ; i += blockDim.x) {
        const int g = bcmap(bq + i, N, bc);
        V8<double> q = zero8<double, MASK>();
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (V::live(c)) q.c[c] = stage_x(a, V::slot(c) * P + off + g, xdt);
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (V::live(c)) qd[V::slot(c)][i] = q.c[c];
        if constexpr (HO) {
            const V8<double> w = c2p<MASK, SPLIT>(q, gamma);
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (V::live(c)) cq[V::slot(c)][i] = w.c[c];
        }
    }
    __syncthreads();
    // B: primitive cell averages wS on the window (FP64, then cast)
    for (int i = threadIdx.x; i < NW; i += blockDim.x) {
        const int iq = e(bw + i) - bq;
        V8<double> w;
        if constexpr (HO) {
            auto Q = [&](int j) {
                V8<double> v = zero8<double, MASK>();
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (V::live(c)) v.c[c] = qd[V::slot(c)][j];
                return v;
            };
            auto CQ = [&](int j) {
                V8<double> v = zero8<double, MASK>();
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (V::live(c)) v.c[c] = cq[V::slot(c)][j];
                return v;
            };
            const V8<double> z8 = zero8<double, MASK>();
            w = ho_cell_prim<MASK, 1, SPLIT>(Q(iq), Q(iq - 1), Q(iq + 1), z8, z8, CQ(iq), CQ(iq - 1), CQ(iq + 1), z8,
                                             z8, gamma);
        } else {
            V8<double> q = zero8<double, MASK>();
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (V::live(c)) q.c[c] = qd[V::slot(c)][iq];
            w = c2p<MASK, SPLIT>(q, gamma);
        }
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (V::live(c)) ws[V::slot(c)][i] = R(w.c[c]);
    }
    __syncthreads();
    struct WAcc {
        R (*ws)[NW];
        int bw, N, bc;
        __device__ int e(int p) const { return bc == EDGE ? clampi(p, 0, N - 1) : p; }
        __device__ V8<R> operator()(int m) const {
            const int j = e(m) - bw;
            V8<R> v = zero8<R, MASK>();
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (VS<MASK>::live(c)) v.c[c] = ws[VS<MASK>::slot(c)][j];
            return v;
        }
    } W{ws, bw, N, bc};
    if constexpr (RECON == PPM && AUTH != AUTH_MC) {
        if (a.flags_only) {
            // the reference's grid-wide interface-limiter tests of this stage
            // (limiters.py:58) over the tile's own cells
            if (threadIdx.x == 0) abits = 0u;
            __syncthreads();
            unsigned bits = 0;
            for (int i = threadIdx.x; i < TS; i += blockDim.x) {
                const int ce = s0 + i;
                if (ce < N) bits |= author_face_tests<R, MASK, AUTH>(W, ce);
            }
            if (bits) atomicOr(&abits, bits);
            __syncthreads();
            if (threadIdx.x == 0 && abits) atomicOr(&ps->aflags, (unsigned long long)abits << (2 * a.stage_idx));
            __syncthreads();  // abits and the LDS windows are reused by the next tile
            continue;
        }
    }
    // C: reconstruction at cells [s0-1, s0+TS+2); their +-4 stencils cover
    // the window [s0-5, s0+TS+6), so tiles with that window inside the domain
    // (and every periodic tile) skip the boundary map
    if constexpr (RECON != PCM) {
        const bool clampC = bc == EDGE && !(s0 >= HW_L && s0 + TS + HW_R <= N);
        bool anyR = false, anyL = false;
        if constexpr (AUTH != AUTH_MC) {
            const unsigned long long af = ld_const(&ps->aflags) >> (2 * a.stage_idx);
            anyR = (af & 1ull) != 0;
            anyL = (af & 2ull) != 0;
        }
        auto recon_cells = [&](auto clamp_tag) {
            constexpr bool CL = decltype(clamp_tag)::value;
            const LineAcc<R, MASK, NW, CL> Wc{ws, bw, N};
            if constexpr (RECON == PPM && AUTH != AUTH_MC) {
                for (int i = threadIdx.x; i < NC; i += blockDim.x) {
                    V8<R> L, Rr;
                    author_cell<R, MASK, AUTH>(Wc, Wc.e(bcl + i), anyR, anyL, L, Rr);
#pragma unroll
                    for (int c = 0; c < 8; ++c)
                        if (V::live(c)) {
                            wl[V::slot(c)][i] = L.c[c];
                            wr[V::slot(c)][i] = Rr.c[c];
                        }
                }
                return;
            }
            if constexpr (RECON == PPM) {
                // the cells' face etas read chi~ at me-1, me, me+1 with
                // me = e(ce-1), e(ce): one evaluation per position
                if (a.flattener != FLAT_OFF) {
                    for (int i = threadIdx.x; i < NCHI; i += blockDim.x)
                        chi_s[i] = flat_chi_at<R, 0>(Wc, a.flattener, s0 - 3 + i);
                    __syncthreads();
                }
            }
            for (int i = threadIdx.x; i < NC; i += blockDim.x) {
                V8<R> L, Rr;
                if constexpr (RECON == PPM)
                    recon_cell<R, MASK, RECON, 0>(Wc, Wc.e(bcl + i), a.limiter, a.flattener, L, Rr,
                                                  [&](int mm) { return chi_s[mm - (s0 - 3)]; });
                else
                    recon_cell<R, MASK, RECON, 0>(Wc, Wc.e(bcl + i), a.limiter, a.flattener, L, Rr);
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (V::live(c)) {
                        wl[V::slot(c)][i] = L.c[c];
                        wr[V::slot(c)][i] = Rr.c[c];
                    }
            }
        };
        if (clampC) recon_cells(std::true_type{});
        else recon_cells(std::false_type{});
        __syncthreads();
    }
    auto WL = [&](int m) {  // wL at cell m (stored values already follow the rule)
        if constexpr (RECON == PCM) return W(m);
        else {
            V8<R> v = zero8<R, MASK>();
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (V::live(c)) v.c[c] = wl[V::slot(c)][m - bcl];
            return v;
        }
    };
    auto WR = [&](int m) {
        if constexpr (RECON == PCM) return W(m);
        else {
            V8<R> v = zero8<R, MASK>();
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (V::live(c)) v.c[c] = wr[V::slot(c)][m - bcl];
            return v;
        }
    };
    // D: spectral radius of the Jacobian state at padded positions j in [s0, s0+TS+1]
    double lmax = 0.0;
    bool bad = false;
    for (int i = threadIdx.x; i < NR; i += blockDim.x) {
        const int j = s0 + i;
        V8<R> st;
        if constexpr (RECON == PCM) {
            st = W(e(j - 1));  // pcm.py:28-30: Jacobian at padded cells
        } else {
            const int je = bc == EDGE ? clampi(j, 1, N) : j;
            const V8<R> wp = WL(je), wm = WR(je - 1);
            if constexpr (SPLIT) st = jac_state<MASK, true, RECON != PLM>(wp, wm);
            else st = (RECON == PLM) ? arith_avg<R, MASK>(wp, wm) : roe_avg<R, MASK>(wp, wm);
        }
        R r;
        if constexpr (SPLIT) r = radius<MASK, 0, true>(st, gamma);
        else r = spectral_radius<R, MASK, 0>(st, gR);
        if constexpr (SOLVER == LW) rad[i] = lw_coef<R>(st, r, gR);
        else rad[i] = r;
        if (j >= s0 + 1 && j <= s0 + TS && j >= 1 && j <= N) {
            const double rd = double(r);
            if (!(rd <= 1.7976931348623157e308)) bad = true;
            lmax = (rd > lmax || rd != rd) ? rd : lmax;
        }
    }
    if (bad) atomicOr(&ps->fail_flag, 1);
    if (a.stage0) block_max_atomic(lmax, &ps->eig[0], red);
    __syncthreads();
    // E: interface fluxes k in [s0, s0+TS]
    for (int i = threadIdx.x; i < NF; i += blockDim.x) {
        const int k = s0 + i;
        Iface<R> f;
        f.wm = WR(k - 1);
        f.wp = WL(k);
        if constexpr (RECON == PCM) {
            auto Qc = [&](int m) {
                V8<double> v = zero8<double, MASK>();
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (V::live(c)) v.c[c] = qd[V::slot(c)][e(m) - bq];
                return cast8<R, double, MASK>(v);
            };
            f.qm = Qc(k - 1);
            f.qp = Qc(k);
            f.wm = W(e(k - 1));
            f.wp = W(e(k));
        } else if constexpr (SPLIT) {
            f.qm = prim_to_cons_split<MASK>(f.wm, gamma);
            f.qp = prim_to_cons_split<MASK>(f.wp, gamma);
        } else {
            f.qm = prim_to_cons<R, MASK>(f.wm, gR);
            f.qp = prim_to_cons<R, MASK>(f.wp, gR);
        }
        if constexpr (SPLIT) {
            f.fm = phys_flux_split<MASK, 0>(f.wm, gamma);
            f.fp = phys_flux_split<MASK, 0>(f.wp, gamma);
        } else {
            f.fm = phys_flux<R, MASK, 0>(f.wm, gR);
            f.fp = phys_flux<R, MASK, 0>(f.wp, gR);
        }
        const R lam = nmax(rad[i], rad[i + 1]);
        R Bn = R(0);
        if constexpr (SOLVER == HLLD) Bn = W(e(k)).c[5];
        V8<R> F = riemann<R, MASK, SOLVER, 0>(f, lam, sax, Bn, gR, 1, a.compat != 0, SPLIT);
        dots_fail<SOLVER, MASK>(F, ps);
        if constexpr (RECON == PPM) {
            if (a.av) {
                // artificial viscosity at interface k (schemes/ppm.py:138-170,
                // McCorquodale & Colella 2011 eq. 35-38): lambda = u_k - u_{k-1}
                // of the cell averages, nu = min(1, (dx lambda)^2 / (beta c_min^2))
                // lambda where lambda < 0 (else 0), c_min the smaller sound speed
                // sqrt(gamma p / rho) of the two cells, alpha = beta = 0.3; the
                // flux gains alpha nu (q_k - q_{k-1}), the diffusive sign
                const V8<R> am = W(e(k - 1)), ap = W(e(k));
                const R lam_av = ap.c[1] - am.c[1];
                if (!(lam_av >= R(0))) {
                    const R cm = ah_sqrt(gdiv(gR * am.c[4], am.c[0])), cp = ah_sqrt(gdiv(gR * ap.c[4], ap.c[0]));
                    const R cmin = nmin(cm, cp);
                    const R dl = R(pd.dx) * lam_av;
                    const R nu = nmin(R(1), gdiv(dl * dl, R(0.3) * (cmin * cmin))) * lam_av;
#pragma unroll
                    for (int c = 0; c < 8; ++c)
                        if (V::live(c)) {
                            const R dq = R(qd[V::slot(c)][e(k) - bq] - qd[V::slot(c)][e(k - 1) - bq]);
                            F.c[c] = F.c[c] + R(0.3) * (nu * dq);
                        }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (V::live(c)) fl[V::slot(c)][i] = F.c[c];
        if (a.flux_out && k <= N && (k < s0 + TS || k == N)) {
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (V::live(c)) a.flux_out[V::slot(c) * a.flux_plane + k] = double(F.c[c]);
        }
    }
    __syncthreads();
    // F: L = -(F[i+1] - F[i]) / dx (evolvers.py:41-60)
    const DivDx bydx(pd.dx);
    int hit_;
    const double fdt = a.fuse_k2 ? step_dt(pd, *ps, 1, &hit_) : 0.0;
    for (int i = threadIdx.x; i < TS; i += blockDim.x) {
        const int cell = s0 + i;
        if (cell >= N) break;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R d = bydx(R(fl[v][i + 1] - fl[v][i]));
            const double L = -double(d);
            const long long j = v * P + off + cell;
            if (a.fuse_k2) {
#pragma clang fp contract(off)
                const double k1 = a.xl ? stage_x(a, j, fdt) : a.fuse_k1[j];
                a.fuse_k2[j] = .25 * (3 * a.fuse_u[j] + k1 + fdt * L);
            } else {
                a.out[j] = L;
            }
        }
    }
    __syncthreads();  // the next tile reuses the LDS windows
    }
}

// Window accessor along the sweep for one line of a 2D tile: W(m) is the
// primitive state at sweep position e(m).  CLAMP = false when the tile's
// stencil (+-4 cells) never leaves the domain or the boundary is periodic
// (the window then holds the wrapped cells): e is the identity and repeated
// positions are recognised as such by the compiler.
template <class R, int MASK, int TT, int NWS, bool CLAMP>
struct RowAcc {
    R (*row)[TT][NWS];
    int it, bws, n;
    __device__ int e(int p) const {
        if constexpr (CLAMP) return clampi(p, 0, n - 1);
        else return p;
    }
    __device__ V8<R> operator()(int m) const {
        const int j = e(m) - bws;
        V8<R> v = zero8<R, MASK>();
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (VS<MASK>::live(c)) v.c[c] = row[VS<MASK>::slot(c)][it][j];
        return v;
    }
};

// ============================================================================
// 2D reconstruction kernel: stage input -> wS (4th-order or point conversion)
// -> limited interface states wL/wR of the tile's own cells, written in the
// sweep-frame layout [t][s] (s contiguous) of type R.
// ============================================================================
// Hydrodynamic tiles take 39 KB of LDS, room for four workgroups per CU; the
// register budget is capped so that four waves fit each SIMD (the PPM
// instantiations spill 80 bytes per lane to do so and still run 10 % faster
// than at three waves on the 1024^2 shear layer).  MHD windows leave LDS for
// three, so their registers stay uncapped.
// AUTH (PPM, 2D hydrodynamics, FP64): author 'c' or 'ph' (ppm.py:43-71).  A
// first launch with flags_only records the stage's grid-wide interface-limiter
// tests of this sweep (limiters.py:58) in the problem's aflags, bits
// 2 (2 stage + AX) and 2 (2 stage + AX) + 1; the second reconstructs with them.
template <class R, int MASK, int RECON, int AX, int TS, int TT, bool SPLIT = false, int AUTH = AUTH_MC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VS<MASK>::NV * (int)sizeof(R) > 32 ? 1 : 4)))
void recon2d_kernel(SweepArgs a) {
    using V = VS<MASK>;
    constexpr int NV = V::NV;
    constexpr bool HO = (RECON == PPM || RECON >= WENO3);
    constexpr int HS = 4;                    // wS window [s0-4, s0+TS+4)
    constexpr int NWS = TS + 2 * HS;
    constexpr int NQS = NWS + 2;             // q window  [s0-5, s0+TS+5)
    constexpr int HT = HO ? 1 : 0;           // q rows    [t0-HT, t0+TT+HT)
    constexpr int NQT = TT + 2 * HT;
    __shared__ double qd[NV][NQT][NQS];
    __shared__ double cq[HO ? NV : 1][HO ? NQT : 1][HO ? NQS : 1];
    __shared__ R ws[NV][TT][NWS];
    // PPM flattener: chi~ of each own row at positions [s0-2, s0+TS]
    constexpr int NCHI = TS + 3;
    __shared__ R chi_s[RECON == PPM ? TT : 1][RECON == PPM ? NCHI : 1];
    constexpr bool AU = RECON == PPM && AUTH != AUTH_MC;
    __shared__ unsigned abits;  // authors c/ph: the tile's interface-limiter tests

    const int ntiles_ = wave_uniform(a.ntiles_dev ? *a.ntiles_dev : a.ntiles);
    for (int ti_ = blockIdx.x; ti_ < ntiles_; ti_ += gridDim.x) {
    const Tile tile = ld_const(a.tiles + ti_);
    const ProbDesc pd = ld_const(a.pd + tile.prob);
    ProbState* ps = a.ps + tile.prob;
    if (!ld_const(&ps->active)) continue;
    if (AU && threadIdx.x == 0) abits = 0u;  // read after the barriers of stages A and B
    const int ns = AX == 0 ? pd.n0 : pd.n1, nt = AX == 0 ? pd.n1 : pd.n0;
    const int bc = pd.bc, s0 = tile.s0, t0 = tile.t0;
    const double gamma = pd.gamma;
    const long long off = pd.off, P = a.plane;
    auto gidx = [&](int s, int t) { return AX == 0 ? s * pd.n1 + t : t * pd.n1 + s; };
    auto es = [&](int p) { return bc == EDGE ? clampi(p, 0, ns - 1) : p; };
    auto et = [&](int p) { return bc == EDGE ? clampi(p, 0, nt - 1) : p; };
    const int bqs = s0 - HS - 1, bqt = t0 - HT, bws = s0 - HS;
    int xhit_;
    const double xdt = a.xl ? step_dt(pd, *ps, 2, &xhit_) : 0.0;

    // A: load, memory-contiguous index fastest
    for (int i = threadIdx.x; i < NQT * NQS; i += blockDim.x) {
        int is, it;
        if constexpr (AX == 0) { is = i / NQT; it = i % NQT; }
        else { it = i / NQS; is = i % NQS; }
        const int gs = bcmap(bqs + is, ns, bc), gt = bcmap(bqt + it, nt, bc);
        const long long g = off + gidx(gs, gt);
        V8<double> q = zero8<double, MASK>();
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (V::live(c)) q.c[c] = stage_x(a, V::slot(c) * P + g, xdt);
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (V::live(c)) qd[V::slot(c)][it][is] = q.c[c];
        if constexpr (HO) {
            const V8<double> w = c2p<MASK, SPLIT>(q, gamma);
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (V::live(c)) cq[V::slot(c)][it][is] = w.c[c];
        }
    }
    __syncthreads();
    // B: wS for s in [s0-4, s0+TS+4), own rows t in [t0, t0+TT)
    for (int i = threadIdx.x; i < TT * NWS; i += blockDim.x) {
        const int it = i / NWS, is = i % NWS;
        const int t = t0 + it;
        const int js = es(bws + is) - bqs;
        const int jt = et(t) - bqt;
        auto Q = [&](int x, int y) {
            V8<double> v = zero8<double, MASK>();
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (V::live(c)) v.c[c] = qd[V::slot(c)][y][x];
            return v;
        };
        V8<double> w;
        if constexpr (HO) {
            auto CQ = [&](int x, int y) {
                V8<double> v = zero8<double, MASK>();
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (V::live(c)) v.c[c] = cq[V::slot(c)][y][x];
                return v;
            };
            w = ho_cell_prim<MASK, 2, SPLIT>(Q(js, jt), Q(js - 1, jt), Q(js + 1, jt), Q(js, jt - 1), Q(js, jt + 1),
                                             CQ(js, jt), CQ(js - 1, jt), CQ(js + 1, jt), CQ(js, jt - 1), CQ(js, jt + 1),
                                             gamma);
        } else {
            w = c2p<MASK, SPLIT>(Q(js, jt), gamma);
        }
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (V::live(c)) ws[V::slot(c)][it][is] = R(w.c[c]);
    }
    __syncthreads();
    // C: reconstruct own cells, write wL/wR (sweep frame [t][s])
    R* wLo = reinterpret_cast<R*>(a.wL);
    R* wRo = reinterpret_cast<R*>(a.wR);
    const long long fo = off;  // same packing as the state, sweep-frame order
    // edge tiles whose +-4 stencil reaches past the domain evaluate the
    // boundary rule; all other tiles (and every periodic tile) do not
    const bool clampS = bc == EDGE && !(s0 >= HS && s0 + TS + HS <= ns);
    auto cells = [&](auto clamp_tag) {
    constexpr bool CL = decltype(clamp_tag)::value;
    if constexpr (AU) {
        const int sh = 2 * (2 * a.stage_idx + AX);
        if (a.flags_only) {
            unsigned bits = 0;
            for (int i = threadIdx.x; i < TT * TS; i += blockDim.x) {
                const int it = i / TS, s = s0 + i % TS, t = t0 + it;
                if (s >= ns || t >= nt) continue;
                const RowAcc<R, MASK, TT, NWS, CL> W{ws, it, bws, ns};
                bits |= author_face_tests<R, MASK, AUTH>(W, s);
            }
            if (bits) atomicOr(&abits, bits);
            __syncthreads();
            if (threadIdx.x == 0 && abits) atomicOr(&ps->aflags, (unsigned long long)abits << sh);
            return;
        }
        const unsigned long long af = ld_const(&ps->aflags) >> sh;
        const bool anyR = (af & 1ull) != 0, anyL = (af & 2ull) != 0;
        for (int i = threadIdx.x; i < TT * TS; i += blockDim.x) {
            const int it = i / TS, s = s0 + i % TS, t = t0 + it;
            if (s >= ns || t >= nt) continue;
            const RowAcc<R, MASK, TT, NWS, CL> W{ws, it, bws, ns};
            V8<R> L, Rr;
            author_cell<R, MASK, AUTH>(W, s, anyR, anyL, L, Rr);
            const long long o = fo + (long long)t * ns + s;
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (V::live(c)) {
                    wLo[V::slot(c) * P + o] = L.c[c];
                    wRo[V::slot(c) * P + o] = Rr.c[c];
                }
        }
        return;
    }
    if constexpr (RECON == PPM) {
        // each cell's face etas read chi~ at me-1, me, me+1 (me = e(ce-1),
        // e(ce)): one evaluation per position instead of up to four per cell
        if (a.flattener != FLAT_OFF) {
            for (int i = threadIdx.x; i < TT * NCHI; i += blockDim.x) {
                const int it = i / NCHI, j = i % NCHI;
                const RowAcc<R, MASK, TT, NWS, CL> W{ws, it, bws, ns};
                chi_s[it][j] = flat_chi_at<R, AX>(W, a.flattener, s0 - 2 + j);
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < TT * TS; i += blockDim.x) {
        const int it = i / TS, is = i % TS;
        const int s = s0 + is, t = t0 + it;
        if (s >= ns || t >= nt) continue;
        const RowAcc<R, MASK, TT, NWS, CL> W{ws, it, bws, ns};
        V8<R> L, Rr;
        if constexpr (RECON == PPM)
            recon_cell<R, MASK, RECON, AX>(W, s, a.limiter, a.flattener, L, Rr,
                                           [&](int mm) { return chi_s[it][mm - (s0 - 2)]; });
        else
            recon_cell<R, MASK, RECON, AX>(W, s, a.limiter, a.flattener, L, Rr);
        const long long o = fo + (long long)t * ns + s;
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (V::live(c)) {
                wLo[V::slot(c) * P + o] = L.c[c];
                wRo[V::slot(c) * P + o] = Rr.c[c];
            }
        if constexpr (V::MHD) {
            // face values of this sweep for the constrained transport: the
            // unlimited interpolant for PPM (ppm.py:38,101), wR otherwise
            // (plm.py:97, weno.py:186; wR = wS for PCM, pcm.py:34)
            if (a.wF) {
                V8<R> f = Rr;
                if constexpr (RECON == PPM) {
                    const V8<R> am = W(s - 1), a0 = W(s), a1 = W(s + 1), a2 = W(s + 2);
#pragma unroll
                    for (int c = 0; c < 8; ++c)
                        if (V::live(c)) f.c[c] = R(K7_12) * (a0.c[c] + a1.c[c]) - R(K1_12) * (am.c[c] + a2.c[c]);
                }
                R* wFo = reinterpret_cast<R*>(a.wF);
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (V::live(c)) wFo[V::slot(c) * P + o] = f.c[c];
            }
            // HLLD's normal field: component (solver axis + 5) of the centred
            // state, the solver axis being the sweep's enumeration index
            // (solvers.py:160,168,183)
            if (a.bn) {
                const int parity = a.parity_override >= 0 ? a.parity_override : ld_const(&ps->parity);
                const int sax = a.compat ? AX : (parity == 0 ? AX : 1 - AX);
                const V8<R> z = W(s);
                reinterpret_cast<R*>(a.bn)[o] = sax == 0 ? z.c[5] : z.c[6];
            }
        }
    }
    };
    if (clampS) cells(std::true_type{});
    else cells(std::false_type{});
    __syncthreads();  // the next tile reuses the LDS windows
    }
}

// ============================================================================
// 2D flux kernel: interface states (sweep frame) -> face conversions ->
// Riemann on face averages -> second Riemann on face-centred states ->
// transverse correction (functions/fv.py:147-153) -> flux difference.
// CONV: 0 = PCM (q from the cell state), 1 = point (PLM), 2 = 4th-order face
// conversion (PPM/WENO, fv.py:105-122 with 'face').
// FRECON >= 0 (hydrodynamics): one launch per sweep.  The tile reconstructs
// (scheme FRECON) the interface states of its own window, the cells
// [s0-1, s0+TS+1] on the lines [t0-2, t0+TT+2), straight into LDS instead of
// reading the wL/wR planes a reconstruction launch wrote: the same cell
// evaluation as recon2d_kernel on the same boundary-mapped window, so the
// window holds bit for bit the states that launch writes.
// ============================================================================
template <class R, int MASK, int SOLVER, int CONV, int AX, int TS, int TT, bool SPLIT = false, int NTHR = 256,
          int FRECON = -1>
__global__ __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(FRECON >= 0 ? 3 : 1)))
void flux2d_kernel(SweepArgs a) {
    using V = VS<MASK>;
    constexpr int NV = V::NV;
    constexpr int NK = TS + 2;   // interfaces [s0, s0+TS+1]
    constexpr int NT = TT + 4;   // w lines   [t0-2, t0+TT+2)
    constexpr int NKF = TS + 1;  // flux interfaces [s0, s0+TS]
    constexpr int NTA = TT + 2;  // face-average lines [t0-1, t0+TT+1)
    constexpr bool FUSED = FRECON >= 0;
    static_assert(!FUSED || !V::MHD, "the in-tile reconstruction writes no constrained-transport face values");
    __shared__ R wm_s[NV][NT][NK];
    __shared__ R wp_s[NV][NT][NK];
    // In-tile reconstruction windows (FUSED): q and its primitive form on the
    // lines [t0-2-FHT, t0+TT+2+FHT) and positions [s0-1-FHS-1, s0+TS+2+FHS],
    // wS on the lines [t0-2, t0+TT+2) and positions [s0-1-FHS, s0+TS+2+FHS),
    // the flattener's chi~ at [s0-3, s0+TS+3).  They are dead once the
    // interface states are in wm_s/wp_s and share their LDS with the
    // face-average arrays of the solve stages.
    constexpr bool FHO = FRECON == PPM || FRECON >= WENO3;
    constexpr int FHS = FHO ? 4 : 1;
    constexpr int FNC = TS + 3;
    constexpr int FNWS = FNC + 2 * FHS;
    constexpr int FNQS = FNWS + 2;
    constexpr int FHT = FHO ? 1 : 0;
    constexpr int FNQT = NT + 2 * FHT;
    constexpr int FNCHI = TS + 6;
    constexpr long long FQ_B = (long long)NV * FNQT * FNQS * 8;
    constexpr long long FREC_B = FQ_B * (FHO ? 2 : 1) + (long long)NV * NT * FNWS * sizeof(R) +
                                 (FRECON == PPM ? (long long)NT * FNCHI * sizeof(R) : 0);
    constexpr long long FFACE_B = 5ll * NV * NTA * NKF * sizeof(R);
    constexpr long long FBUF_B = FUSED ? (FREC_B > FFACE_B ? FREC_B : FFACE_B) : 8;
    __shared__ double fbuf[(FBUF_B + 7) / 8];
    constexpr int FA = FUSED ? 1 : NV, FB = FUSED ? 1 : NTA, FC = FUSED ? 1 : NKF;
    __shared__ R qm_a[FA][FB][FC];
    __shared__ R qp_a[FA][FB][FC];
    __shared__ R fm_a[FA][FB][FC];
    __shared__ R fp_a[FA][FB][FC];
    __shared__ R fa_a[FA][FB][FC];
    using FaceArr = R (*)[NTA][NKF];
    auto face_arr = [&](R* own, int k) -> FaceArr {
        if constexpr (FUSED)
            return reinterpret_cast<FaceArr>(reinterpret_cast<char*>(fbuf) + k * (long long)NV * NTA * NKF * sizeof(R));
        else
            return reinterpret_cast<FaceArr>(own);
    };
    const FaceArr qm_s = face_arr(&qm_a[0][0][0], 0), qp_s = face_arr(&qp_a[0][0][0], 1),
                  fm_s = face_arr(&fm_a[0][0][0], 2), fp_s = face_arr(&fp_a[0][0][0], 3),
                  fa_s = face_arr(&fa_a[0][0][0], 4);
    // only LLF and Lax-Wendroff read coefficients of the halo positions (the
    // radii, or ss^2 / radius); the other solvers need the radius of the
    // tile's own interfaces (dt reduction and the non-finite check), which
    // stage B then evaluates alone
    constexpr bool LAM = SOLVER == LLF || SOLVER == LW;
    __shared__ R rad_s[LAM ? NTA : 1][LAM ? NK : 1];
    __shared__ R bn_s[SOLVER == HLLD ? NTA : 1][SOLVER == HLLD ? NKF : 1];
    // When every thread owns at most one interface of stage D, the corrected
    // fluxes wait in registers across a barrier and then take the place of
    // qm_s (dead by then), which saves the F window's LDS: the 16x10 hydro
    // tile then needs 50.5 KB and three workgroups fit a CU instead of two.
    constexpr bool FALIAS = TT * NKF <= NTHR;
    __shared__ R F_own[FALIAS ? 1 : NV][FALIAS ? 1 : TT][FALIAS ? 1 : NKF];
    R (*F_s)[TT][NKF] = reinterpret_cast<R (*)[TT][NKF]>(FALIAS ? &qm_s[0][0][0] : &F_own[0][0][0]);
    __shared__ double red[NTHR / 64];

    const int ntiles_ = wave_uniform(a.ntiles_dev ? *a.ntiles_dev : a.ntiles);
    const R* wLg = reinterpret_cast<const R*>(a.wL);
    const R* wRg = reinterpret_cast<const R*>(a.wR);
    const long long P = a.plane;
    // w-/w+ window of a tile, interfaces [s0, s0+TS+1] x lines [t0-2, t0+TT+2),
    // fetched into registers one tile ahead: the loads of the next tile are in
    // flight while this tile computes (residency is bounded by the LDS windows
    // before the registers, so the prefetch costs no occupancy)
    constexpr int NA = FUSED ? 1 : (NT * NK + NTHR - 1) / NTHR;
    R pfm[NA][NV], pfp[NA][NV];
    auto fetch = [&](int ti) {
        if constexpr (FUSED) return;
        const Tile tl = ld_const(a.tiles + ti);
        const ProbDesc q = ld_const(a.pd + tl.prob);
        const int ns_ = AX == 0 ? q.n0 : q.n1, nt_ = AX == 0 ? q.n1 : q.n0;
#pragma unroll
        for (int j = 0; j < NA; ++j) {
            const int i = threadIdx.x + j * NTHR;
            if (i < NT * NK) {
                const int it = i / NK, ik = i % NK;
                const int k = tl.s0 + ik;
                const long long row = q.off + (long long)bcmap(tl.t0 - 2 + it, nt_, q.bc) * ns_;
                const long long om = row + bcmap(k - 1, ns_, q.bc), op = row + bcmap(k, ns_, q.bc);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    pfm[j][v] = wRg[v * P + om];
                    pfp[j][v] = wLg[v * P + op];
                }
            }
        }
    };
    if ((int)blockIdx.x < ntiles_) fetch(blockIdx.x);
    for (int ti_ = blockIdx.x; ti_ < ntiles_; ti_ += gridDim.x) {
    const Tile tile = ld_const(a.tiles + ti_);
    const ProbDesc pd = ld_const(a.pd + tile.prob);
    ProbState* ps = a.ps + tile.prob;
    const bool active = ld_const(&ps->active);
    // A: the prefetched window into LDS, then start the next tile's loads
    if (!FUSED && active) {
#pragma unroll
        for (int j = 0; j < NA; ++j) {
            const int i = threadIdx.x + j * NTHR;
            if (i < NT * NK) {
                const int it = i / NK, ik = i % NK;
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    wm_s[v][it][ik] = pfm[j][v];
                    wp_s[v][it][ik] = pfp[j][v];
                }
            }
        }
    }
    if (ti_ + (int)gridDim.x < ntiles_) fetch(ti_ + gridDim.x);
    if (!active) continue;
    const int ns = AX == 0 ? pd.n0 : pd.n1, nt = AX == 0 ? pd.n1 : pd.n0;
    const int bc = pd.bc, s0 = tile.s0, t0 = tile.t0;
    const R gR = R(pd.gamma);
    const long long off = pd.off;
    const int parity = a.parity_override >= 0 ? a.parity_override : ld_const(&ps->parity);
    // solver axis: the enumeration index of the sweep (solvers.py:34-36), or
    // the physical axis in the physical-compat mode
    const int sax = a.compat ? AX : (parity == 0 ? AX : 1 - AX);
    const int bt = t0 - 2, bta = t0 - 1;
    auto sidx = [&](int s, int t) { return off + (long long)t * ns + s; };  // sweep-frame layout
    auto gidx = [&](int s, int t) { return off + (AX == 0 ? (long long)s * pd.n1 + t : (long long)t * pd.n1 + s); };
    if constexpr (FUSED) {
        // R-A: q (and its primitive form) of the window, boundary mapped
        using QArr = double (*)[FNQT][FNQS];
        using WArr = R (*)[NT][FNWS];
        const QArr qd = reinterpret_cast<QArr>(fbuf);
        const QArr cq = reinterpret_cast<QArr>(reinterpret_cast<char*>(fbuf) + FQ_B);
        const WArr ws = reinterpret_cast<WArr>(reinterpret_cast<char*>(fbuf) + FQ_B * (FHO ? 2 : 1));
        R (*chi_s)[FNCHI] =
            reinterpret_cast<R (*)[FNCHI]>(reinterpret_cast<char*>(ws) + (long long)NV * NT * FNWS * sizeof(R));
        const double gamma = pd.gamma;
        const int bqs = s0 - 1 - FHS - 1, bqt = t0 - 2 - FHT, bws = s0 - 1 - FHS;
        auto es = [&](int p) { return bc == EDGE ? clampi(p, 0, ns - 1) : p; };
        auto etl = [&](int p) { return bc == EDGE ? clampi(p, 0, nt - 1) : p; };
        int rhit_;
        const double rxdt = a.xl ? step_dt(pd, *ps, 2, &rhit_) : 0.0;
        for (int i = threadIdx.x; i < FNQT * FNQS; i += blockDim.x) {
            int is, it;
            if constexpr (AX == 0) { is = i / FNQT; it = i % FNQT; }
            else { it = i / FNQS; is = i % FNQS; }
            const int gs = bcmap(bqs + is, ns, bc), gt = bcmap(bqt + it, nt, bc);
            const long long g = gidx(gs, gt);
            V8<double> q = zero8<double, MASK>();
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (V::live(c)) q.c[c] = stage_x(a, V::slot(c) * P + g, rxdt);
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (V::live(c)) qd[V::slot(c)][it][is] = q.c[c];
            if constexpr (FHO) {
                const V8<double> w = c2p<MASK, SPLIT>(q, gamma);
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (V::live(c)) cq[V::slot(c)][it][is] = w.c[c];
            }
        }
        __syncthreads();
        // R-B: wS on every line of the window (line t0-2+ir evaluated on the
        // boundary-mapped line, as the reconstruction launch evaluates its own)
        for (int i = threadIdx.x; i < NT * FNWS; i += blockDim.x) {
            const int ir = i / FNWS, is = i % FNWS;
            const int js = es(bws + is) - bqs;
            const int jt = etl(t0 - 2 + ir) - bqt;
            auto Q = [&](int x, int y) {
                V8<double> v = zero8<double, MASK>();
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (V::live(c)) v.c[c] = qd[V::slot(c)][y][x];
                return v;
            };
            V8<double> w;
            if constexpr (FHO) {
                auto CQ = [&](int x, int y) {
                    V8<double> v = zero8<double, MASK>();
#pragma unroll
                    for (int c = 0; c < 8; ++c)
                        if (V::live(c)) v.c[c] = cq[V::slot(c)][y][x];
                    return v;
                };
                w = ho_cell_prim<MASK, 2, SPLIT>(Q(js, jt), Q(js - 1, jt), Q(js + 1, jt), Q(js, jt - 1),
                                                 Q(js, jt + 1), CQ(js, jt), CQ(js - 1, jt), CQ(js + 1, jt),
                                                 CQ(js, jt - 1), CQ(js, jt + 1), gamma);
            } else {
                w = c2p<MASK, SPLIT>(Q(js, jt), gamma);
            }
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (V::live(c)) ws[V::slot(c)][ir][is] = R(w.c[c]);
        }
        __syncthreads();
        // R-C: cells [s0-1, s0+TS+1] of every line: cell s0-1+j gives w- of
        // interface s0+j (its wR) and w+ of interface s0+j-1 (its wL)
        const bool clampS = bc == EDGE && !(s0 - 1 - FHS >= 0 && s0 + TS + 1 + FHS < ns);
        auto rcells = [&](auto clamp_tag) {
            constexpr bool CL = decltype(clamp_tag)::value;
            if constexpr (FRECON == PPM) {
                if (a.flattener != FLAT_OFF) {
                    for (int i = threadIdx.x; i < NT * FNCHI; i += blockDim.x) {
                        const int ir = i / FNCHI, j = i % FNCHI;
                        const RowAcc<R, MASK, NT, FNWS, CL> W{ws, ir, bws, ns};
                        chi_s[ir][j] = flat_chi_at<R, AX>(W, a.flattener, s0 - 3 + j);
                    }
                    __syncthreads();
                }
            }
            for (int i = threadIdx.x; i < NT * FNC; i += blockDim.x) {
                const int ir = i / FNC, j = i % FNC;
                const int v = s0 - 1 + j;
                const int ce = CL ? clampi(v, 0, ns - 1) : v;
                const RowAcc<R, MASK, NT, FNWS, CL> W{ws, ir, bws, ns};
                V8<R> L, Rr;
                if constexpr (FRECON == PPM)
                    recon_cell<R, MASK, FRECON, AX>(W, ce, a.limiter, a.flattener, L, Rr,
                                                    [&](int mm) { return chi_s[ir][mm - (s0 - 3)]; });
                else
                    recon_cell<R, MASK, FRECON, AX>(W, ce, a.limiter, a.flattener, L, Rr);
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (V::live(c)) {
                        if (j < NK) wm_s[V::slot(c)][ir][j] = Rr.c[c];
                        if (j >= 1) wp_s[V::slot(c)][ir][j - 1] = L.c[c];
                    }
            }
        };
        if (clampS) rcells(std::true_type{});
        else rcells(std::false_type{});
    }
    __syncthreads();
    auto WM = [&](int it, int ik) {
        V8<R> w = zero8<R, MASK>();
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (V::live(c)) w.c[c] = wm_s[V::slot(c)][it][ik];
        return w;
    };
    auto WP = [&](int it, int ik) {
        V8<R> w = zero8<R, MASK>();
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (V::live(c)) w.c[c] = wp_s[V::slot(c)][it][ik];
        return w;
    };
    // Tiles whose lines [t0-2, t0+TT+2) and Jacobian positions [s0, s0+TS+1]
    // stay inside the domain (and every periodic tile) skip the boundary map.
    const bool clampT = bc == EDGE && !(t0 >= 2 && t0 + TT + 2 <= nt && s0 >= 1 && s0 + TS + 1 <= ns);
    R Freg[NV];
    auto stages = [&](auto clamp_tag) {
    constexpr bool CL = decltype(clamp_tag)::value;
    auto et = [&](int p) {
        if constexpr (CL) return clampi(p, 0, nt - 1);
        else return p;
    };
    // B: spectral radius at padded Jacobian positions j in [s0, s0+TS+1],
    // lines [t0-1, t0+TT+1) (evaluated on the boundary-mapped line); without
    // LLF only at the own positions j in [s0+1, s0+TS], lines [t0, t0+TT)
    double lmax = 0.0;
    bool bad = false;
    constexpr int NB = LAM ? NTA * NK : TT * TS;
    for (int i = threadIdx.x; i < NB; i += blockDim.x) {
        const int ia = LAM ? i / NK : 1 + i / TS, ik = LAM ? i % NK : 1 + i % TS;
        const int t = bta + ia, te = et(t), it = te - bt;
        const int j = s0 + ik;
        V8<R> st;
        if constexpr (CONV == 0) {
            // PCM: Jacobian at cell j-1 = w_minus at interface j
            st = WM(it, ik);
        } else {
            const int je = CL ? clampi(j, 1, ns) : j;
            const int jk = je - s0;  // in [0, NK) or the clamped edge interface
            const V8<R> wp = WP(it, jk), wm = WM(it, jk);
            if constexpr (SPLIT) st = jac_state<MASK, true, CONV != 1>(wp, wm);
            else st = (CONV == 1) ? arith_avg<R, MASK>(wp, wm) : roe_avg<R, MASK>(wp, wm);
        }
        R r;
        if constexpr (SPLIT) r = radius<MASK, AX, true>(st, pd.gamma);
        else r = spectral_radius<R, MASK, AX>(st, gR);
        if constexpr (SOLVER == LW) rad_s[ia][ik] = lw_coef<R>(st, r, gR);
        else if constexpr (LAM) rad_s[ia][ik] = r;
        if (t >= t0 && t < t0 + TT && t < nt && j >= s0 + 1 && j <= s0 + TS && j >= 1 && j <= ns) {
            const double rd = double(r);
            if (!(rd <= 1.7976931348623157e308)) bad = true;
            lmax = (rd > lmax || rd != rd) ? rd : lmax;
        }
    }
    if (bad) atomicOr(&ps->fail_flag, 1);
    if (a.stage0) block_max_atomic(lmax, &ps->eig[AX], red);
    __syncthreads();
    // C: q+-, f+- and the face-average flux on lines [t0-1, t0+TT+1)
    int xhit_;
    const double xdt = (CONV == 0 && a.xl) ? step_dt(pd, *ps, 2, &xhit_) : 0.0;
    for (int i = threadIdx.x; i < NTA * NKF; i += blockDim.x) {
        const int ia = i / NKF, ik = i % NKF;
        const int t = bta + ia, te = et(t), it = te - bt;
        const int k = s0 + ik;
        Iface<R> f;
        f.wm = WM(it, ik);
        f.wp = WP(it, ik);
        if constexpr (CONV == 0) {
            const int gt = bcmap(te, nt, bc);
            V8<double> qm = zero8<double, MASK>(), qp = zero8<double, MASK>();
            const long long om = gidx(bcmap(k - 1, ns, bc), gt), op = gidx(bcmap(k, ns, bc), gt);
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (V::live(c)) {
                    qm.c[c] = stage_x(a, V::slot(c) * P + om, xdt);
                    qp.c[c] = stage_x(a, V::slot(c) * P + op, xdt);
                }
            f.qm = cast8<R, double, MASK>(qm);
            f.qp = cast8<R, double, MASK>(qp);
        } else if constexpr (CONV == 1) {
            f.qm = p2c<MASK, SPLIT>(f.wm, pd.gamma);
            f.qp = p2c<MASK, SPLIT>(f.wp, pd.gamma);
        } else {
            // face conversion along t (fv.py:105-122, axes = [1] of the sweep frame):
            // q = P(w - K24 lap_t w) + K24 lap_t P(w)
            auto face = [&](const V8<R>& wc, const V8<R>& wlo, const V8<R>& whi) {
                const V8<R> pc = p2c<MASK, SPLIT>(wc, pd.gamma), plo = p2c<MASK, SPLIT>(wlo, pd.gamma),
                            phi = p2c<MASK, SPLIT>(whi, pd.gamma);
                V8<R> ww = wc;
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (V::live(c)) ww.c[c] = wc.c[c] - R(K24) * ((whi.c[c] - wc.c[c]) - (wc.c[c] - wlo.c[c]));
                V8<R> q = p2c<MASK, SPLIT>(ww, pd.gamma);
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (V::live(c)) q.c[c] = q.c[c] + R(K24) * ((phi.c[c] - pc.c[c]) - (pc.c[c] - plo.c[c]));
                return q;
            };
            const int itm = et(te - 1) - bt, itp = et(te + 1) - bt;
            f.qm = face(f.wm, WM(itm, ik), WM(itp, ik));
            f.qp = face(f.wp, WP(itm, ik), WP(itp, ik));
        }
        f.fm = pflux<MASK, AX, SPLIT>(f.wm, pd.gamma);
        f.fp = pflux<MASK, AX, SPLIT>(f.wp, pd.gamma);
        R Bn = R(0);
        if constexpr (SOLVER == HLLD) {
            const int gt = bcmap(te, nt, bc);
            Bn = reinterpret_cast<const R*>(a.bn)[sidx(bcmap(k, ns, bc), gt)];
            bn_s[ia][ik] = Bn;
        }
        const R lam = LAM ? nmax(rad_s[ia][ik], rad_s[ia][ik + 1]) : R(0);
        const V8<R> F = riemann<R, MASK, SOLVER, AX>(f, lam, sax, Bn, gR, sax + 1, a.compat != 0, SPLIT);
        dots_fail<SOLVER, MASK>(F, ps);
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (V::live(c)) {
                const int v = V::slot(c);
                qm_s[v][ia][ik] = f.qm.c[c];
                qp_s[v][ia][ik] = f.qp.c[c];
                fm_s[v][ia][ik] = f.fm.c[c];
                fp_s[v][ia][ik] = f.fp.c[c];
                fa_s[v][ia][ik] = F.c[c];
            }
    }
    __syncthreads();
    // D: centred second solve + transverse correction on own lines
    for (int i = threadIdx.x; i < TT * NKF; i += blockDim.x) {
        const int itt = i / NKF, ik = i % NKF;
        const int t = t0 + itt;
        const int te = et(t);
        const int ia0 = te - bta, iam = et(te - 1) - bta, iap = et(te + 1) - bta;
        const int it0 = te - bt, itm = et(te - 1) - bt, itp = et(te + 1) - bt;
        auto corr = [&](R (*arr)[NT][NK], int v) {
            const R x0 = arr[v][it0][ik], xm = arr[v][itm][ik], xp = arr[v][itp][ik];
            return x0 - R(K24) * ((xp - x0) - (x0 - xm));
        };
        auto corrA = [&](R (*arr)[NTA][NKF], int v) {
            const R x0 = arr[v][ia0][ik], xm = arr[v][iam][ik], xp = arr[v][iap][ik];
            return x0 - R(K24) * ((xp - x0) - (x0 - xm));
        };
        Iface<R> f;
        f.wm = zero8<R, MASK>();
        f.wp = f.qm = f.qp = f.fm = f.fp = f.wm;
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (V::live(c)) {
                const int v = V::slot(c);
                f.wm.c[c] = corr(wm_s, v);
                f.wp.c[c] = corr(wp_s, v);
                f.qm.c[c] = corrA(qm_s, v);
                f.qp.c[c] = corrA(qp_s, v);
                f.fm.c[c] = corrA(fm_s, v);
                f.fp.c[c] = corrA(fp_s, v);
            }
        const R lam = LAM ? nmax(rad_s[ia0][ik], rad_s[ia0][ik + 1]) : R(0);
        R Bn = R(0);
        if constexpr (SOLVER == HLLD) Bn = bn_s[ia0][ik];
        const V8<R> Fc = riemann<R, MASK, SOLVER, AX>(f, lam, sax, Bn, gR, sax + 1, a.compat != 0, SPLIT);
        dots_fail<SOLVER, MASK>(Fc, ps);
        const int k = s0 + ik;
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (V::live(c)) {
                const int v = V::slot(c);
                const R F = Fc.c[c] - R(K24) * ((fa_s[v][iap][ik] - fa_s[v][ia0][ik]) - (fa_s[v][ia0][ik] - fa_s[v][iam][ik]));
                if constexpr (FALIAS) Freg[v] = F;
                else F_s[v][itt][ik] = F;
                if (a.flux_out && t < nt && k <= ns && (k < s0 + TS || k == ns))
                    a.flux_out[v * a.flux_plane + (long long)k * nt + t] = double(F);
            }
    }
        };
    if (clampT) stages(std::true_type{});
    else stages(std::false_type{});
    __syncthreads();
    if constexpr (FALIAS) {
        // interface i = threadIdx.x of stage D (one per thread at most)
        const int i = threadIdx.x;
        if (i < TT * NKF) {
#pragma unroll
            for (int v = 0; v < NV; ++v) F_s[v][i / NKF][i % NKF] = Freg[v];
        }
        __syncthreads();
    }
    // E: flux difference; first sweep writes D, second writes L = -(D0 + D1)
    const DivDx bydx(pd.dx);
    int hit_;
    const double fdt = a.fuse_k2 ? step_dt(pd, *ps, 2, &hit_) : 0.0;
    for (int i = threadIdx.x; i < TT * TS; i += blockDim.x) {
        int itt, is;
        if constexpr (AX == 0) { is = i / TT; itt = i % TT; }
        else { itt = i / TS; is = i % TS; }
        const int s = s0 + is, t = t0 + itt;
        if (s >= ns || t >= nt) continue;
        const long long g = gidx(s, t);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double d = double(bydx(R(F_s[v][itt][is + 1] - F_s[v][itt][is])));
            if (a.last) {
                const double L = -(a.Dprev[v * P + g] + d);
                if (a.fuse_k2) {
#pragma clang fp contract(off)
                    const long long j = v * P + g;
                    const double k1 = a.xl ? stage_x(a, j, fdt) : a.fuse_k1[j];
                    a.fuse_k2[j] = .25 * (3 * a.fuse_u[j] + k1 + fdt * L);
                } else {
                    a.out[v * P + g] = L;
                }
            } else {
                a.out[v * P + g] = d;
            }
        }
    }
    __syncthreads();  // the next tile reuses the LDS windows
    }
}

}  // namespace ah

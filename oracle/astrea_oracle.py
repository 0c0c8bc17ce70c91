"""CPU oracle for astrea's finite-volume update -- TEST INFRASTRUCTURE ONLY.

This module is a from-scratch numpy restatement of the reference algorithm
(mervyzr/astrea 1.3.2).  It exists to check the HIP path and to time a CPU
baseline: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg may import it, and never as the thing measured or shipped.

Every function cites the reference file:line it restates.  Floating-point
operations are written in the same order as the reference so that, with
``eig="closed"``, results are reproducible to the last bit by a kernel that
evaluates the same expressions; ``eig="lapack"`` keeps the reference's cost
structure (an 8x8 eigenvalue solve per interface, functions/fv.py:157-169).

Conventions (sweep frame): every sweep works on the state transposed so the
sweep axis is axis 0; the variable axis (last, 8 entries: rho, vx, vy, vz,
p|E, Bx, By, Bz) is never permuted.  Interface k in [0, N] separates cells
k-1 and k; ``w_minus[k]`` is the right face of cell k-1 and ``w_plus[k]`` the
left face of cell k.  Every intermediate array that the reference pads is
evaluated at the clamped (edge) or wrapped (wrap) index of that array.

Parity status per function is listed in ``PARITY``; tests/test_oracle_golden.py
checks every pinned entry against tests/golden/*.npz written from the
reference itself, or against closed forms and invariants.
"""
from dataclasses import dataclass, field, replace
import itertools

import numpy as np

# Python float literals the reference uses; kept as named constants so the
# rounding of each coefficient is the same as in the reference expressions.
C24 = 1 / 24
C7_12, C1_12 = 7 / 12, 1 / 12

PARITY = {
    "point/high-order conversions": "closed form + golden (initialise.npz)",
    "make_flux": "closed form (tests) + golden via schemes_1d fplus/fminus",
    "roe_average": "closed form + golden via steps",
    "limiters.slope (minmod..superbee)": "textbook phi(r) closed forms",
    "extrapolant_limiter mc": "golden limiter_mc.npz",
    "interface_limiter": "golden pins.npz iflim_* (reference limiters.interface_limiter: data with and without "
                         "local extrema)",
    "ppm authors c / ph": "golden pins.npz ppm_c_*, ppm_ph_* (reference ppm.run with author 'c'/'ph', "
                          "hydro and MHD, both boundaries)",
    "Osher-Solomon (DOTS) flux + right eigenvectors": "golden schemes_1d.npz *_flux_os (bitwise, 24 cases)",
    "pcm/plm/ppm/weno interface states": "golden schemes_1d.npz",
    "spectral radius (closed)": "golden localmax (LAPACK) to 1e-13 rel",
    "llf/hllc/hlld fluxes": "golden schemes_1d.npz",
    "es flux": "golden schemes_1d.npz *_flux_es (NaN positions included)",
    "lw flux (flux_lw, LAPACK pick)": "golden schemes_1d.npz (12 hydro cases)",
    "lw flux (sound-speed form)": "structure pinned through flux_lw_speed + consistency; "
                                  "the speed choice departs from the reference by design",
    "2D double solve + transverse correction": "golden space_2d.npz",
    "MHD constrained transport": "golden space_2d.npz + steps ot16_*",
    "SSPRK/RK stages": "golden steps rk_* + amplification polynomials",
    "driver loop / dt": "golden steps dts",
    "ppm flattener (astrea)": "golden pins.npz flat_* (reference apply_flattener called in isolation, "
                              "1D/2D, both axes and boundaries, exact switch points) + eta=1 invariant",
    "ppm flattener (colella90)": "Colella's shock test: smooth data unflattened, eta<1 only where it fires",
    "hllc compat=physical": "golden pins.npz mirror_* (y sweep = reference x sweep of the mirrored grid), "
                            "space_2d x sweep, consistency at axis 1, upwind supersonic limits, "
                            "transposition equivariance, convergence to the exact Riemann solution",
    "hlld compat=physical": "golden space_2d p0 (the reference's call at the physical axis)",
    "artificial viscosity (mu)": "golden pins.npz av_strong_* (reference apply_artificial_viscosity, the "
                                 "N = 8 case it runs on, saturated coefficient) + known answer of eq. 36",
    "artificial viscosity (flux hook)": "build option, off by default (the reference never applies mu)",
}


# ----------------------------------------------------------------------------
# Options (the hot-path fields of astrea's sim_variables; generic.py:159-286)
# ----------------------------------------------------------------------------
@dataclass
class Scheme:
    dimension: int = 1
    subgrid: str = "plm"          # pcm | plm | ppm | weno3 | weno5 | weno7
    solver: str = "llf"           # llf | lw | hllc | hlld | es | os
    timestep: str = "ssprk(3,3)"  # euler | rk4 | ssprk(2,2) ... ssprk(10,4)
    boundary: str = "edge"        # edge (outlet) | wrap (periodic)
    gamma: float = 1.4
    cfl: float = .5
    dx: float = 1 / 128
    magnetic_2d: bool = False
    limiter: str = "minmod"       # PLM slope limiter (reference hard-codes minmod)
    flattener: str = "off"        # off | astrea | colella90 (PPM only)
    eig: str = "closed"           # closed | lapack
    parity: int = 0               # step parity: sweep order flips every step
    compat: str = "astrea"        # astrea: reference quirks | physical: solver axis = sweep axis
    category: str = ""            # solver_category when it differs from the name's (generic.py:241,266-269)
    av: bool = False              # artificial viscosity on the fluxes (1D PPM; the reference never applies it)
    ppm_author: str = "mc"        # mc | c | ph (ppm.py:13; evolvers.py:17 always passes 'mc')

    @property
    def high_order(self):
        return self.subgrid.startswith("w") or self.subgrid in ("ppm", "parabolic", "p")

    @property
    def solver_category(self):
        """generic.py:241 derives the category from the solver name; a
        magnetic configuration with HLLC then has its name reset to 'lf' while
        the category stays 'hll' (generic.py:266-269), which ``category``
        carries."""
        return self.category or {"llf": "lax", "lf": "lax", "lw": "lax", "hllc": "hll", "hlld": "hll",
                                  "os": "complete", "es": "complete"}[self.solver]

    def order(self):
        """Sweep order: physical axes in the order the reference iterates its
        permutations dict, which reverses every step (astrea.py:85)."""
        axes = list(range(self.dimension))
        return axes if self.parity == 0 else axes[::-1]


def perm(axis, dim):
    """generic.py:258-263: sweep-frame transposition for physical axis."""
    ps = [a for a in itertools.permutations(range(dim + 1)) if a[-1] == dim]
    return ps[axis]


# ----------------------------------------------------------------------------
# Elementary operations (functions/fv.py)
# ----------------------------------------------------------------------------
def gdiv(a, b):
    """fv.py:19-20: a/b where b != 0, else 0."""
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    out = np.zeros(np.broadcast(a, b).shape)
    np.divide(a, b, out=out, where=(b != 0))
    return out


def glog(a):
    """fv.py:25-26: log where a > 0, else 0."""
    out = np.zeros_like(a)
    np.log(a, out=out, where=a > 0)
    return out


def sum3(x):
    """numpy add.reduce over a length-3 last axis: ((x0 + x1) + x2)."""
    return (x[..., 0] + x[..., 1]) + x[..., 2]


def norm2(v):
    """fv.norm(v)**2 as the reference computes it (fv.py:37-38): the square
    of the rounded Euclidean norm."""
    n = np.sqrt(sum3(v * v))
    return n * n


def pad(x, bc, width=1, axis=0):
    """fv.py:57-61 (np.pad with mode edge|wrap along one axis)."""
    p = [(0, 0)] * x.ndim
    p[axis] = (width, width)
    return np.pad(x, p, mode=bc)


def lap(x, bc, axis):
    """fv.py:42-44 on an array padded by one along ``axis``:
    (x[i+1] - x[i]) - (x[i] - x[i-1])."""
    xp = pad(x, bc, 1, axis)
    n = x.shape[axis]
    up = np.take(xp, range(2, n + 2), axis=axis)
    mid = np.take(xp, range(1, n + 1), axis=axis)
    lo = np.take(xp, range(0, n), axis=axis)
    return (up - mid) - (mid - lo)


def prim_to_cons(w, gamma):
    """fv.py:89-93 with fv.py:49-51."""
    q = np.copy(w)
    q[..., 4] = w[..., 4] / (gamma - 1) + .5 * (w[..., 0] * norm2(w[..., 1:4]) + norm2(w[..., 5:8]))
    q[..., 1:4] = w[..., 1:4] * w[..., 0][..., None]
    return q


def cons_to_prim(q, gamma):
    """fv.py:97-101 with fv.py:52-53."""
    w = np.copy(q)
    v = gdiv(q[..., 1:4], q[..., 0][..., None])
    w[..., 4] = (gamma - 1) * (q[..., 4] - .5 * (q[..., 0] * norm2(v) + norm2(q[..., 5:8])))
    w[..., 1:4] = v
    return w


def ho_convert(kind, x, bc, axes):
    """fv.py:67-85: x -/+ Lap/24 along each axis in ``axes`` (sequentially,
    each Laplacian taken of the unmodified input).  kind 'avg' subtracts."""
    out = np.copy(x)
    for ax in axes:
        if kind == "avg":
            out -= C24 * lap(x, bc, ax)
        else:
            out += C24 * lap(x, bc, ax)
    return out


def ho_prim_to_cons(w, gamma, bc, axes):
    """fv.py:105-122 (4th-order averaged primitive -> conservative)."""
    ww, q = np.copy(w), np.zeros_like(w)
    pw = prim_to_cons(w, gamma)
    for ax in axes:
        ww -= C24 * lap(w, bc, ax)
        q += C24 * lap(pw, bc, ax)
    return prim_to_cons(ww, gamma) + q


def ho_cons_to_prim(q, gamma, bc, axes):
    """fv.py:126-143 (4th-order averaged conservative -> primitive)."""
    w, qq = np.zeros_like(q), np.copy(q)
    cq = cons_to_prim(q, gamma)
    for ax in axes:
        qq -= C24 * lap(q, bc, ax)
        w += C24 * lap(cq, bc, ax)
    return cons_to_prim(qq, gamma) + w


def to_prim(q, s, kind="cell"):
    """sim_variables.convert_conservative (generic.py:250-255)."""
    if s.high_order:
        axes = range(s.dimension) if kind == "cell" else range(1, s.dimension)
        return ho_cons_to_prim(q, s.gamma, s.boundary, axes)
    return cons_to_prim(q, s.gamma)


def to_cons(w, s, kind="cell"):
    """sim_variables.convert_primitive (generic.py:250-255)."""
    if s.high_order:
        axes = range(s.dimension) if kind == "cell" else range(1, s.dimension)
        return ho_prim_to_cons(w, s.gamma, s.boundary, axes)
    return prim_to_cons(w, s.gamma)


# ----------------------------------------------------------------------------
# Fluxes, Roe average, spectral radius (functions/constructor.py, fv.py)
# ----------------------------------------------------------------------------
def phys_flux(w, gamma, axis):
    """constructor.py:113-125: ideal-MHD flux normal to ``axis``."""
    rho, v, p, B = w[..., 0], w[..., 1:4], w[..., 4], w[..., 5:8]
    a, o, c = axis % 3, (axis + 1) % 3, (axis + 2) % 3
    f = np.zeros_like(w)
    b2 = norm2(B)
    f[..., 0] = rho * v[..., axis]
    f[..., a + 1] = rho * v[..., axis] ** 2 + p + .5 * b2 - B[..., axis] ** 2
    f[..., o + 1] = rho * v[..., axis] * v[..., o] - B[..., axis] * B[..., o]
    f[..., c + 1] = rho * v[..., axis] * v[..., c] - B[..., axis] * B[..., c]
    f[..., 4] = v[..., axis] * (.5 * rho * norm2(v) + (gamma * p) / (gamma - 1) + b2) - B[..., axis] * sum3(v * B)
    f[..., o + 5] = B[..., o] * v[..., axis] - B[..., axis] * v[..., o]
    f[..., c + 5] = B[..., c] * v[..., axis] - B[..., axis] * v[..., c]
    return f


def roe_average(wp, wm):
    """constructor.py:167-176 (B components cross-weighted, line 174)."""
    out = np.zeros_like(wp)
    rm, rp = np.sqrt(wm[..., 0]), np.sqrt(wp[..., 0])
    s = rm + rp
    out[..., 0] = rm * rp
    out[..., 1:4] = gdiv(wp[..., 1:4] * rp[..., None] + wm[..., 1:4] * rm[..., None], s[..., None])
    out[..., 4] = gdiv(rp * wp[..., 4] + rm * wm[..., 4], s)
    out[..., 5:8] = gdiv(wp[..., 5:8] * rm[..., None] + wm[..., 5:8] * rp[..., None], s[..., None])
    return out


def jacobian(w, gamma, axis):
    """constructor.py:129-163: 8x8 primitive-variable Jacobian."""
    rho, v, p, B = w[..., 0], w[..., 1:4], w[..., 4], w[..., 5:8]
    a, o, c = axis % 3, (axis + 1) % 3, (axis + 2) % 3
    A = np.zeros(w.shape + (8,))
    for i in range(8):
        A[..., i, i] = v[..., a]
    A[..., 0, a + 1] = rho
    A[..., a + 1, 4] = 1 / rho
    A[..., 4, a + 1] = gamma * p
    A[..., a + 5, a + 5] = 0
    A[..., a + 1, a + 5] = -gdiv(B[..., a], rho)
    A[..., a + 1, o + 5] = gdiv(B[..., o], rho)
    A[..., a + 1, c + 5] = gdiv(B[..., c], rho)
    A[..., o + 1, o + 5] = -gdiv(B[..., a], rho)
    A[..., c + 1, c + 5] = -gdiv(B[..., a], rho)
    A[..., o + 1, a + 5] = -gdiv(B[..., o], rho)
    A[..., c + 1, a + 5] = -gdiv(B[..., c], rho)
    A[..., o + 5, a + 1] = B[..., o]
    A[..., c + 5, a + 1] = B[..., c]
    A[..., 4, a + 5] = (gamma - 1) * sum3(v * B)
    A[..., o + 5, o + 1] = -B[..., a]
    A[..., c + 5, c + 1] = -B[..., a]
    A[..., o + 5, a + 5] = -v[..., o]
    A[..., c + 5, a + 5] = -v[..., c]
    return A


def spectral_radius(w, gamma, axis, mhd):
    """Closed-form max|eigenvalue| of ``jacobian(w)`` (replaces the LAPACK
    call at fv.py:157-161).  The B-normal row of the Jacobian is zero, so one
    eigenvalue is 0 and the rest are those of the 7x7 primitive MHD Jacobian:
    u, u +- c_a, u +- c_s, u +- c_f.  Hence rho(A) = |u| + c_f, with
    c_f^2 = (a^2 + b^2 + sqrt((a^2 + b^2)^2 - 4 a^2 b_n^2)) / 2,
    a^2 = gamma p / rho, b^2 = |B|^2 / rho, b_n^2 = B_n^2 / rho.  For B = 0 this
    is |u| + sqrt(gamma p / rho); a negative a^2 gives the modulus of the
    complex pair, sqrt(u^2 - a^2)."""
    rho, u, p = w[..., 0], w[..., axis + 1], w[..., 4]
    a2 = gamma * p / rho
    if not mhd:
        with np.errstate(invalid="ignore"):
            return np.where(a2 >= 0, np.abs(u) + np.sqrt(np.abs(a2)), np.sqrt(np.abs(u * u - a2)))
    B = w[..., 5:8]
    b2 = sum3(B * B) / rho
    bn2 = B[..., axis] * B[..., axis] / rho
    s = a2 + b2
    cf2 = .5 * (s + np.sqrt(np.maximum(s * s - 4 * a2 * bn2, 0.)))
    return np.abs(u) + np.sqrt(np.maximum(cf2, 0.))


def local_max_eig(w_jac, s, axis):
    """fv.py:157-161: max|lambda| of the Jacobian at each entry."""
    if s.eig == "lapack":
        chars = np.linalg.eigvals(jacobian(w_jac, s.gamma, axis))
        return np.max(np.abs(chars), axis=-1), chars
    mhd = s.magnetic_2d or bool(np.any(w_jac[..., 5:8] != 0))
    return spectral_radius(w_jac, s.gamma, axis, mhd), None


# ----------------------------------------------------------------------------
# Limiters (num_methods/limiters.py)
# ----------------------------------------------------------------------------
def slope_limiter(wp, name):
    """limiters.py:10-49 on an array padded by one along axis 0.
    a = backward difference, b = forward difference."""
    a = wp[1:-1] - wp[:-2]
    b = wp[2:] - wp[1:-1]
    if name == "minmod":
        out = np.zeros_like(b)
        ab = a * b > 0
        m1 = (np.abs(a) < np.abs(b)) & ab
        out[m1] = a[m1]
        m2 = (np.abs(a) >= np.abs(b)) & ab
        out[m2] = b[m2]
        return out
    r = gdiv(a, b)
    if name == "vanleer":
        phi = (r + np.abs(r)) / (1 + np.abs(r))
    elif name == "ospre":
        phi = 1.5 * ((r ** 2 + r) / (r ** 2 + r + 1))
    elif name == "vanalbada1":
        phi = (r ** 2 + r) / (r ** 2 + 1)
    elif name == "koren":
        phi = np.maximum(np.zeros_like(r), np.minimum(np.minimum(2 * r, (2 + r) / 3), np.full_like(r, 2)))
    elif name == "superbee":
        phi = np.maximum(np.zeros_like(r), np.maximum(np.minimum(2 * r, np.ones_like(r)),
                                                      np.minimum(r, np.full_like(r, 2))))
    else:
        raise ValueError(name)
    return phi * b


def mc_extrapolant_limiter(wS, bc, wF_L, wF_R):
    """limiters.py:82-143, author 'mc' (McCorquodale & Colella 2011).

    The reference tests ``cell_extrema.any()`` over the whole grid (line 102)
    and otherwise takes a simpler branch (141-143).  That test is always true:
    with edge ghosts the first cell has w[-2] == w[0], so (w - w[-2])(w[+2] - w)
    is 0; with periodic ghosts the chain i, i+2, i+4, ... returns to its start,
    so its differences cannot all share one sign.  Only the taken branch is
    restated."""
    C = 5 / 4
    w1 = pad(wS, bc, 1)
    w2 = pad(wS, bc, 2)
    wm1, wp1 = w1[:-2], w1[2:]
    wm2, wp2 = w2[:-4], w2[4:]
    dw_m, dw_p = wS - wF_L, wF_R - wS
    wL, wR = np.copy(wF_L), np.copy(wF_R)
    d2w = 6 * (wF_L - 2 * wS + wF_R)
    d2w_C = wm1 - 2 * wS + wp1
    d2p = pad(d2w_C, bc, 1)
    d3w = np.diff(d2p, axis=0)[1:]
    cell_extrema = (dw_m * dw_p <= 0) | ((wS - wm2) * (wp2 - wS) <= 0)
    lim_curv = np.sign(d2w_C) * np.minimum(np.minimum(np.abs(d2w), C * np.abs(d2w_C)),
                                           np.minimum(C * np.abs(d2p[2:]), C * np.abs(d2p[:-2])))
    d2w_lim = np.zeros_like(wS)
    d2w_lim[cell_extrema] = lim_curv[cell_extrema]
    rho_lim = np.zeros_like(wS)
    sens = np.abs(d2w) > 1e-12 * np.maximum(np.abs(wS), np.maximum(np.maximum(np.abs(wm1), np.abs(wp1)),
                                                                    np.maximum(np.abs(wm2), np.abs(wp2))))
    phi = gdiv(d2w_lim, d2w)
    rho_lim[sens] = phi[sens]
    d3p2 = pad(d3w, bc, 2)
    d3p1 = d3p2[1:-1]
    d3min = np.minimum(np.minimum(d3p1[:-2], d3w), np.minimum(d3p2[:-4], d3p2[4:]))
    d3max = np.maximum(np.maximum(d3p1[:-2], d3w), np.maximum(d3p2[:-4], d3p2[4:]))
    roundoff = (rho_lim < (1 - 1e-12)) | (.1 * np.maximum(np.abs(d3max), np.abs(d3min)) <= (d3max - d3min))
    m = (dw_m * dw_p < 0) & roundoff
    wL[m] = (wS - rho_lim * dw_m)[m]
    wR[m] = (wS + rho_lim * dw_p)[m]
    m = roundoff & (np.abs(dw_m) >= 2 * np.abs(dw_p))
    wL[m] = (wS - 2 * (1 - rho_lim) * dw_p - rho_lim * dw_m)[m]
    m = roundoff & (np.abs(dw_p) >= 2 * np.abs(dw_m))
    wR[m] = (wS + 2 * (1 - rho_lim) * dw_m + rho_lim * dw_p)[m]
    return wL, wR


def interface_limiter(wf, wm1, w0, wp1, wp2):
    """limiters.py:53-79 (Colella et al. 2011 eq. 84-87; Peterson & Hammett
    2008 eq. 3.33-3.34) for face values wf between cells w0 and wp1.  The
    reference decides over the whole array at once (line 58): if no face
    anywhere is a local extremum, every face keeps its value; otherwise every
    face is replaced by the average of its two cells minus a sixth of the
    limited curvature, which is non-zero only where the three curvature
    estimates share one sign."""
    if not np.any((wf - w0) * (wp1 - wf) < 0):
        return wf
    C = 5 / 4
    dl = wm1 - 2 * w0 + wp1
    dc = 3 * (w0 - 2 * wf + wp1)
    dr = w0 - 2 * wp1 + wp2
    same = (np.sign(dl) == np.sign(dr)) & (np.sign(dc) == np.sign(dr)) & (np.sign(dc) == np.sign(dl))
    lim = np.sign(dc) * np.minimum(np.abs(dc), np.minimum(np.abs(C * dl), np.abs(C * dr)))
    d2 = np.zeros_like(wf)
    d2[same] = lim[same]
    return .5 * (w0 + wp1) - d2 / 6


def colella_extrapolant_limiter(wS, w, w2, wFp, wF_L, wF_R, author):
    """limiters.py:144-201: the extrapolant limiter of PPM authors 'c'
    (Colella et al. 2011 eq. 89-98) and 'ph' (Peterson & Hammett 2008 eq.
    3.31-3.39).  ``w``/``w2`` are wS padded by 1/2 cells, ``wFp`` the face
    values padded by 2.  The reference's global test (line 163) leaves the
    face values unlimited when no cell or extrapolant anywhere is an
    extremum; otherwise phi = D2_lim / D2 scales every cell, and D2_lim is
    zero away from extrema (those cells become piecewise constant)."""
    C = 5 / 4
    dm, dp = wS - wF_L, wF_R - wS
    cell_ext = dm * dp <= 0
    ph = author == "ph"
    if ph:
        ext_ext = (w[:-2] - wS) * (wS - w[2:]) <= 0
    else:
        dfl, dfr = wF_L - wFp[:-4], wFp[4:] - wF_R
        dsl, dsr = wS - w[:-2], w[2:] - wS
        df = np.minimum(np.abs(dfl), np.abs(dfr))
        ds = np.minimum(np.abs(dsl), np.abs(dsr))
        ext_ext = ((df >= ds) & (dfl * dfr < 0)) | ((ds >= df) & (dsl * dsr < 0))
    if not (np.any(cell_ext) or np.any(ext_ext)):
        return wF_L, wF_R
    d2 = 6 * (wF_L - 2 * wS + wF_R)
    d2l = w2[:-4] - 2 * w[:-2] + wS
    d2c = w[:-2] - 2 * wS + w[2:]
    d2r = wS - 2 * w[2:] + w2[4:]
    sg = [np.sign(x) for x in (d2, d2c, d2l, d2r)]
    same = ((sg[0] == sg[1]) & (sg[0] == sg[2]) & (sg[0] == sg[3]) & (sg[1] == sg[2]) & (sg[1] == sg[3])
            & (sg[2] == sg[3]))
    lim = sg[0] * np.minimum(np.minimum(np.abs(d2), np.abs(C * d2c)), np.minimum(np.abs(C * d2l), np.abs(C * d2r)))
    d2lim = np.zeros_like(wS)
    m = cell_ext & same
    d2lim[m] = lim[m]
    if ph:
        phi = gdiv(d2lim, d2)
        return wS + phi * (wF_L - wS), wS + phi * (wF_R - wS)
    m = ext_ext & same
    d2lim[m] = lim[m]
    phi = gdiv(d2lim, d2)
    dl, dr = np.copy(dm), np.copy(dp)
    m = np.abs(dm) > 2 * np.abs(dp)
    dl[m] = 2 * dp[m]
    m = np.abs(dp) > 2 * np.abs(dm)
    dr[m] = 2 * dm[m]
    return wS - phi * dl, wS + phi * dr


def flattener_eta(wS, axis, bc, variant, delta=.33, z0=.75, z1=.85):
    """PPM slope flattener (schemes/ppm.py:111-134, Colella 1990).

    ``variant='astrea'`` is the reference's apply_flattener (its column 0:
    the reference broadcasts eta over the 8 variables), bitwise equal to it
    on the pins.npz vectors.  The reference's blend at ppm.py:67 cannot run
    (broadcast error for N != 8) and no option reaches it, so the blend in
    ``reconstruct`` applies eta per cell to all variables.  The literal
    condition of ppm.py:124 sets eta~ := 0 where the flow is expanding AND
    the pressure jump is weak; ``'colella90'`` sets eta~ := 1 unless the zone
    is compressive AND the jump strong, i.e. keeps the ratio ramp only where
    Colella's shock test fires."""
    p1, p2 = pad(wS[..., 4], bc, 1), pad(wS[..., 4], bc, 2)
    u1 = pad(wS[..., axis + 1], bc, 1)
    dp1 = p1[2:] - p1[:-2]
    dp2 = p2[4:] - p2[:-4]
    z = gdiv(np.abs(dp1), np.abs(dp2))
    chi = np.copy(1 - gdiv(z - z0, z1 - z0))
    chi[z > z1] = 0
    chi[z < z0] = 1
    weak = gdiv(np.abs(dp1), np.minimum(p1[2:], p1[:-2])) <= delta
    expanding = (u1[:-2] - u1[2:]) <= 0
    if variant == "astrea":
        chi[expanding & weak] = 0
    else:
        chi[expanding | weak] = 1
    cp = pad(chi, bc, 1)
    sg = np.sign(dp1)
    eta = np.copy(chi)
    eta[sg < 0] = np.minimum(chi, cp[2:])[sg < 0]
    eta[sg > 0] = np.minimum(chi, cp[:-2])[sg > 0]
    return eta


def artificial_viscosity(wS, axis, s, alpha=.3, beta=.3):
    """schemes/ppm.py:138-170 (McCorquodale & Colella 2011 eq. 35-38) per
    face i+1/2, i = 0..N-1, in 1D: lambda = u_{i+1} - u_i (padded), c_min the
    smaller of the two cells' sqrt(fv.divide(gamma p, rho)), nu = min(1,
    (dx lambda)^2 / (beta c_min^2)) lambda, zero where lambda >= 0, and
    mu = alpha (nu (w_{i+1} - w_i)).  This is the reference's output where
    its broadcast lets it run (N = 8, where nu[i, j] = coefficient_j lambda_i
    equals the per-face value whenever the coefficient is the same on every
    face, e.g. saturated at 1); its 2D branch adds arrays of different shapes
    (ppm.py:149-154) and never runs.  Nothing in the reference reads mu."""
    if wS.ndim != 2:
        raise NotImplementedError("the reference's 2D artificial viscosity cannot run (ppm.py:149-154)")
    w = pad(wS, s.boundary)
    lam = w[2:, axis + 1] - w[1:-1, axis + 1]
    cs = np.sqrt(gdiv(s.gamma * w[..., 4], w[..., 0]))
    cmin = np.minimum(cs[1:-1], cs[2:])
    nu = np.minimum(1, gdiv((s.dx * lam) ** 2, beta * cmin ** 2)) * lam
    nu[lam >= 0] = 0
    return alpha * (nu[:, None] * np.diff(w[1:], axis=0))


def av_flux(q, wS, axis, s, alpha=.3, beta=.3):
    """The artificial-viscosity flux the device adds when the option is on,
    at every interface k = 0..N (cells k-1 and k, padded): alpha (nu (q_k -
    q_{k-1})) with nu of ``artificial_viscosity`` and the conservative cell
    averages q (M&C 2011 eq. 38 applies nu to the conserved variables; the
    reference's mu takes differences of the primitive ones).  nu <= 0, so
    adding this term to the flux is diffusive."""
    w = pad(wS, s.boundary)
    qq = pad(q, s.boundary)
    lam = w[1:, axis + 1] - w[:-1, axis + 1]
    cs = np.sqrt(gdiv(s.gamma * w[..., 4], w[..., 0]))
    cmin = np.minimum(cs[:-1], cs[1:])
    nu = np.minimum(1, gdiv((s.dx * lam) ** 2, beta * cmin ** 2)) * lam
    nu[lam >= 0] = 0
    return alpha * (nu[:, None] * (qq[1:] - qq[:-1]))


# ----------------------------------------------------------------------------
# Reconstruction (schemes/*.py) -- sweep frame, returns the data dict
# ----------------------------------------------------------------------------
def weno(wS, bc, order):
    """schemes/weno.py:22-149 (component-wise WENO-3/5/7, eps = 1e-6)."""
    eps = 1e-6
    if order == 3:
        w = pad(wS, bc, 1)
        z, m1, p1 = w[1:-1], w[:-2], w[2:]
        b0, b1 = (z - m1) ** 2, (p1 - z) ** 2
        a0 = lambda g: g / (b0 + eps) ** 2
        a1 = lambda g: g / (b1 + eps) ** 2
        g0, g1 = 1 / 3, 2 / 3
        wR = (a0(g0) / (a0(g0) + a1(g1))) * (1.5 * z - .5 * m1) + (a1(g1) / (a0(g0) + a1(g1))) * (.5 * z + .5 * p1)
        wL = (a1(g0) / (a0(g1) + a1(g0))) * (1.5 * z - .5 * p1) + (a0(g1) / (a0(g1) + a1(g0))) * (.5 * z + .5 * m1)
        return wL, wR
    if order == 7:
        w3 = pad(wS, bc, 3)
        w2 = w3[1:-1]
        w = w2[1:-1]
        z = w[1:-1]
        m1, m2, m3 = w[:-2], w2[:-4], w3[:-6]
        p1, p2, p3 = w[2:], w2[4:], w3[6:]
        g0, g1, g2, g3 = 1 / 35, 12 / 35, 18 / 35, 4 / 35
        b0 = (m3 * (547 * m3 - 3882 * m2 + 4642 * m1 - 1854 * z) + m2 * (7043 * m2 - 17246 * m1 + 7042 * z)
              + m1 * (11003 * m1 - 9402 * z) + z * (2107 * z))
        b1 = (m2 * (267 * m2 - 1642 * m1 + 1602 * z - 494 * p1) + m1 * (2843 * m1 - 5966 * z + 1922 * p1)
              + z * (3443 * z - 2522 * p1) + p1 * (547 * p1))
        b2 = (m1 * (547 * m1 - 2522 * z + 1922 * p1 - 494 * p2) + z * (3443 * z - 5966 * p1 + 1602 * p2)
              + p1 * (2843 * p1 - 1642 * p2) + p2 * (267 * p2))
        b3 = (z * (2107 * z - 9402 * p1 + 7042 * p2 - 1854 * p3) + p1 * (11003 * p1 - 17246 * p2 + 4642 * p3)
              + p2 * (7043 * p2 - 3882 * p3) + p3 * (547 * p3))
        A = [lambda g, b=b: g / (b + eps) ** 2 for b in (b0, b1, b2, b3)]
        sR = A[0](g0) + A[1](g1) + A[2](g2) + A[3](g3)
        wR = ((A[0](g0) / sR) * (-1 / 4 * m3 + 13 / 12 * m2 - 23 / 12 * m1 + 25 / 12 * z)
              + (A[1](g1) / sR) * (1 / 12 * m2 - 5 / 12 * m1 + 13 / 12 * z + 1 / 4 * p1)
              + (A[2](g2) / sR) * (-1 / 12 * m1 + 7 / 12 * z + 7 / 12 * p1 - 1 / 12 * p2)
              + (A[3](g3) / sR) * (1 / 4 * z + 13 / 12 * p1 - 5 / 12 * p2 + 1 / 12 * p3))
        sL = A[0](g3) + A[1](g2) + A[2](g1) + A[3](g0)
        wL = ((A[0](g3) / sL) * (1 / 4 * z + 13 / 12 * m1 - 5 / 12 * m2 + 1 / 12 * m3)
              + (A[1](g2) / sL) * (-1 / 12 * p1 + 7 / 12 * z + 7 / 12 * m1 - 1 / 12 * m2)
              + (A[2](g1) / sL) * (1 / 12 * p2 - 5 / 12 * p1 + 13 / 12 * z + 1 / 4 * m1)
              + (A[3](g0) / sL) * (-1 / 4 * p3 + 13 / 12 * p2 - 23 / 12 * p1 + 25 / 12 * z))
        return wL, wR
    w2 = pad(wS, bc, 2)
    w = w2[1:-1]
    z, m1, m2, p1, p2 = w[1:-1], w[:-2], w2[:-4], w[2:], w2[4:]
    g0, g1, g2 = 1 / 10, 3 / 5, 3 / 10
    b0 = 13 / 12 * (m2 - 2 * m1 + z) ** 2 + 1 / 4 * (m2 - 4 * m1 + 3 * z) ** 2
    b1 = 13 / 12 * (m1 - 2 * z + p1) ** 2 + 1 / 4 * (m1 - p1) ** 2
    b2 = 13 / 12 * (z - 2 * p1 + p2) ** 2 + 1 / 4 * (3 * z - 4 * p1 + p2) ** 2
    A = [lambda g, b=b: g / (b + eps) ** 2 for b in (b0, b1, b2)]
    sR = A[0](g0) + A[1](g1) + A[2](g2)
    wR = ((A[0](g0) / sR) * (1 / 3 * m2 - 7 / 6 * m1 + 11 / 6 * z)
          + (A[1](g1) / sR) * (-1 / 6 * m1 + 5 / 6 * z + 1 / 3 * p1)
          + (A[2](g2) / sR) * (1 / 3 * z + 5 / 6 * p1 - 1 / 6 * p2))
    sL = A[0](g2) + A[1](g1) + A[2](g0)
    wL = ((A[0](g2) / sL) * (1 / 3 * z + 5 / 6 * m1 - 1 / 6 * m2)
          + (A[1](g1) / sL) * (-1 / 6 * p1 + 5 / 6 * z + 1 / 3 * m1)
          + (A[2](g0) / sL) * (1 / 3 * p2 - 7 / 6 * p1 + 11 / 6 * z))
    return wL, wR


def weno_order(subgrid):
    """weno.py:159-165: order from the subgrid string, bare 'weno' is 5."""
    tail = subgrid.replace('-', '').replace(' ', '').split("weno")[-1]
    return int(tail) if tail in ("3", "5", "7") else 5


def reconstruct(q, s, axis):
    """One sweep of schemes/{pcm,plm,ppm,weno}.run on the sweep-frame
    conservative state ``q``.  Returns the per-axis data dict."""
    bc, gamma = s.boundary, s.gamma
    wS = to_prim(q, s, "cell")
    d = {"wS": wS}
    if s.subgrid in ("pcm", "constant", "c", "polynomial"):
        # pcm.py:10-39 (the DB's 'polynomial' also lands here, evolvers.py:20-21)
        w = pad(wS, bc)
        qq = pad(q, bc)
        d.update(wF=wS, wFs=(w[1:], w[:-1]), qFs=(qq[1:], qq[:-1]),
                 jac_state=w, wLR=(wS, wS))
        d["fluxFs"] = (phys_flux(w[1:], gamma, axis), phys_flux(w[:-1], gamma, axis))
        return d
    if s.subgrid in ("plm", "linear", "l"):
        # plm.py:12-62
        grad = .5 * slope_limiter(pad(wS, bc), s.limiter)
        wL, wR = np.copy(wS - grad), np.copy(wS + grad)
        wF = wR
    elif s.subgrid in ("ppm", "parabolic", "p"):
        # ppm.py:13-107, author 'mc'
        w2 = pad(wS, bc, 2)
        w = w2[1:-1]
        wF = C7_12 * (wS + w[2:]) - C1_12 * (w[:-2] + w2[4:])
        wF_unlim = wF
        if s.ppm_author == "ph":
            # ppm.py:43-57: both faces of each cell interpolated and limited separately
            wF_L = C7_12 * (w[:-2] + wS) - C1_12 * (w2[:-4] + w[2:])
            lim = (interface_limiter(wF_L, w2[:-4], w[:-2], wS, w[2:]),
                   interface_limiter(wF, w[:-2], wS, w[2:], w2[4:]))
            wL, wR = colella_extrapolant_limiter(wS, w, w2, None, lim[0], lim[1], "ph")
        elif s.ppm_author == "c":
            # ppm.py:60-62, 69-71: the face values limited, then the extrapolants
            wF = interface_limiter(wF, w[:-2], wS, w[2:], w2[4:])
            wFp = pad(wF, bc, 2)
            wL, wR = colella_extrapolant_limiter(wS, w, w2, wFp, np.copy(wFp[1:-3]), np.copy(wFp[2:-2]), "c")
        else:
            if s.flattener != "off":
                eta = flattener_eta(wS, axis, bc, s.flattener)
                wF = wF * eta[..., None] + wS * (1 - eta)[..., None]
            wFp = pad(wF, bc, 2)
            wL, wR = mc_extrapolant_limiter(wS, bc, np.copy(wFp[1:-3]), np.copy(wFp[2:-2]))
        d["wT_src"] = wF_unlim     # ppm.py:40-41 reads wF before the blend
    else:
        wL, wR = weno(wS, bc, weno_order(s.subgrid))
        wF = wR
    w_plus, w_minus = pad(wL, bc)[1:], pad(wR, bc)[:-1]
    if s.subgrid in ("plm", "linear", "l"):
        avg = (.5 * (w_plus + w_minus))[1:]
        qp, qm = prim_to_cons(w_plus, gamma), prim_to_cons(w_minus, gamma)
    else:
        avg = roe_average(w_plus, w_minus)[1:]
        qp, qm = to_cons(w_plus, s, "face"), to_cons(w_minus, s, "face")
    d.update(wF=wF, wFs=(w_plus, w_minus), qFs=(qp, qm), jac_state=pad(avg, bc), wLR=(wL, wR))
    d["fluxFs"] = (phys_flux(w_plus, gamma, axis), phys_flux(w_minus, gamma, axis))
    return d


# ----------------------------------------------------------------------------
# Riemann solvers (num_methods/solvers.py) -- sweep frame
# ----------------------------------------------------------------------------
def flux_llf(lmax, d):
    """solvers.py:69-75; lmax = per-entry max|lambda| on the padded Jacobian
    states, paired (k, k+1) as at lines 73-74."""
    qp, qm = d["qFs"]
    fp, fm = d["fluxFs"]
    lam = np.maximum(lmax[:-1], lmax[1:])
    return .5 * (fm + fp) - .5 * ((qp - qm) * lam[..., None])


def flux_lw_speed(ss, lmax, d):
    """solvers.py:82-88 for a per-entry wave speed ``ss`` and max|lambda|
    ``lmax`` on the padded Jacobian states: ss^2 / max|lambda| (guarded),
    paired (k, k+1) as LLF pairs its radii, as the dissipation coefficient."""
    qp, qm = d["qFs"]
    fp, fm = d["fluxFs"]
    ne = gdiv(ss ** 2, lmax)
    mx = np.max([ne[:-1], ne[1:]], axis=0)
    return .5 * (fm + fp) - .5 * ((qp - qm) * mx[..., None])


def lw_sound_speed(w_jac, gamma):
    """The wave speed solvers.py:83 names ("sound speed for each cell"),
    sqrt(gamma p / rho) at the Jacobian states.  The closed-form path uses it
    in place of the reference's column pick (``flux_lw``), which cannot be
    reproduced without LAPACK: parity with the reference is not claimed."""
    return np.sqrt(gamma * w_jac[..., 4] / w_jac[..., 0])


def flux_lw(chars, d):
    """solvers.py:79-88 as written: the 'sound speed' is the second unique
    eigenvalue column, np.unique(chars, axis=-1)[..., 1] -- whole columns of
    the flattened LAPACK output sorted lexicographically, so the pick is a
    global decision that depends on LAPACK's eigenvalue order at every entry
    (the zero eigenvalue's column varies with the state)."""
    return flux_lw_speed(np.unique(chars, axis=-1)[..., 1], np.max(np.abs(chars), axis=-1), d)


def flux_hllc(ax, gamma, d, mom=1, default_minus=False):
    """solvers.py:92-138 with ``ax`` the solver's axis argument (the
    enumeration index of the sweep, solvers.py:34-36).  Keeps the reference's
    star momentum at index 1 (127-128) and the F(w+) default (133)."""
    wp, wm = d["wFs"]
    QR, QL = d["qFs"]
    fp, fm = d["fluxFs"]
    rhoL, uL, pL = wm[..., 0], wm[..., ax + 1], wm[..., 4]
    rhoR, uR, pR = wp[..., 0], wp[..., ax + 1], wp[..., 4]
    cL, cR = np.sqrt(gamma * gdiv(pL, rhoL)), np.sqrt(gamma * gdiv(pR, rhoR))
    sqL, sqR = np.sqrt(rhoL), np.sqrt(rhoR)
    u_hat = gdiv(uL * sqL + uR * sqR, sqL + sqR)
    c2_hat = gdiv(sqL * cL ** 2 + sqR * cR ** 2, sqL + sqR) + .5 * ((uR - uL) ** 2) * gdiv(sqL * sqR, (sqL + sqR) ** 2)
    sL = np.minimum(uL - cL, u_hat - np.sqrt(c2_hat))
    sR = np.maximum(uR + cR, u_hat + np.sqrt(c2_hat))
    sM = gdiv(pR - pL + rhoL * uL * (sL - uL) - rhoR * uR * (sR - uR), rhoL * (sL - uL) - rhoR * (sR - uR))
    kL, kR = gdiv(sL - uL, sL - sM), gdiv(sR - uR, sR - sM)
    QLs, QRs = QL * kL[..., None], QR * kR[..., None]
    QLs[..., mom] = rhoL * kL * sM
    QRs[..., mom] = rhoR * kR * sM
    QLs[..., 4] = QLs[..., 4] + kL * (sM - uL) * (rhoL * sM + gdiv(pL, sL - uL))
    QRs[..., 4] = QRs[..., 4] + kR * (sM - uR) * (rhoR * sM + gdiv(pR, sR - uR))
    F = np.copy(fm if default_minus else fp)
    FLs, FRs = fm + (QLs - QL) * sL[..., None], fp + (QRs - QR) * sR[..., None]
    m = (sL <= 0) & (0 < sM)
    F[m] = FLs[m]
    m = (sM <= 0) & (0 <= sR)
    F[m] = FRs[m]
    m = sR < 0
    F[m] = fp[m]
    return F


def flux_hlld(ax, gamma, bc, d):
    """solvers.py:142-232 (Miyoshi & Kusano 2005 as written: fast speed uses
    B[...,0] as the normal field, s_R = min(u) + max(c_f), both Alfven speeds
    s_M - B_n/sqrt(rho*), B_n taken from the right cell's centred state)."""
    wp, wm = d["wFs"]
    QR, QL = d["qFs"]
    fp, fm = d["fluxFs"]
    a, o, c = ax % 3, (ax + 1) % 3, (ax + 2) % 3

    def cfast(w):
        rho, p, B = w[..., 0], w[..., 4], w[..., 5:8]
        cs = np.sqrt(gdiv(gamma * p, rho))
        ca = gdiv(np.sqrt(sum3(B * B)), np.sqrt(rho))
        cax = gdiv(B[..., 0], np.sqrt(rho))
        return np.sqrt(.5 * (cs ** 2 + ca ** 2 + np.sqrt(((cs ** 2 + ca ** 2) ** 2) - (4 * (cs ** 2) * (cax ** 2)))))

    wS = pad(d["wS"], bc)[1:]
    Bn = wS[..., a + 5]
    rhoL, vL, pL, BL = wm[..., 0], wm[..., 1:4], wm[..., 4], wm[..., 5:8]
    rhoR, vR, pR, BR = wp[..., 0], wp[..., 1:4], wp[..., 4], wp[..., 5:8]
    cfL, cfR = cfast(wm), cfast(wp)
    uL, uR = vL[..., ax], vR[..., ax]
    sL = np.minimum(uL, uR) - np.maximum(cfL, cfR)
    sR = np.minimum(uL, uR) + np.maximum(cfL, cfR)
    bL2, bR2 = norm2(BL), norm2(BR)
    sM = gdiv(pR - pL + rhoL * uL * (sL - uL) - rhoR * uR * (sR - uR) + .5 * bR2 - .5 * bL2,
              rhoL * (sL - uL) - rhoR * (sR - uR))
    rLs, rRs = rhoL * gdiv(sL - uL, sL - sM), rhoR * gdiv(sR - uR, sR - sM)
    sLs, sRs = sM - gdiv(BL[..., ax], np.sqrt(rLs)), sM - gdiv(BR[..., ax], np.sqrt(rRs))
    ps = gdiv(rhoL * (pR + .5 * bR2) * (sL - uL) - rhoR * (pL + .5 * bL2) * (sR - uR)
              + rhoR * rhoL * (sL - uL) * (sR - uR), rhoL * (sL - uL) - rhoR * (sR - uR))
    denL = rhoL * (sL - uL) * (sL - sM) - Bn ** 2
    denR = rhoR * (sR - uR) * (sR - sM) - Bn ** 2
    vyLs = vL[..., o] - Bn * BL[..., o] * gdiv(sM - uL, denL)
    vyRs = vR[..., o] - Bn * BR[..., o] * gdiv(sM - uR, denR)
    vzLs = vL[..., c] - Bn * BL[..., c] * gdiv(sM - uL, denL)
    vzRs = vR[..., c] - Bn * BR[..., c] * gdiv(sM - uR, denR)
    ByLs = BL[..., o] * gdiv(rhoL * (sL - uL) ** 2 - Bn ** 2, denL)
    ByRs = BR[..., o] * gdiv(rhoR * (sR - uR) ** 2 - Bn ** 2, denR)
    BzLs = BL[..., c] * gdiv(rhoL * (sL - uL) ** 2 - Bn ** 2, denL)
    BzRs = BR[..., c] * gdiv(rhoR * (sR - uR) ** 2 - Bn ** 2, denR)

    def star(Q, rho, rs, vy, vz, By, Bz):
        S = np.zeros_like(Q)
        S[..., 0] = rs
        S[..., a + 1] = rho * sM
        S[..., o + 1] = rho * vy
        S[..., c + 1] = rho * vz
        S[..., a + 5] = np.copy(Q[..., a + 5])
        S[..., o + 5] = By
        S[..., c + 5] = Bz
        return S

    QLs = star(QL, rhoL, rLs, vyLs, vzLs, ByLs, BzLs)
    QRs = star(QR, rhoR, rRs, vyRs, vzRs, ByRs, BzRs)
    QLs[..., 4] = gdiv(QL[..., 4] * (sL - uL) - uL * (pL + .5 * bL2) + ps * sM
                       + Bn * (sum3(vL * BL) - sum3(QLs[..., 1:4] * QLs[..., 5:8])), sL - sM)
    QRs[..., 4] = gdiv(QR[..., 4] * (sR - uR) - uR * (pR + .5 * bR2) + ps * sM
                       + Bn * (sum3(vR * BR) - sum3(QRs[..., 1:4] * QRs[..., 5:8])), sR - sM)
    FLs, FRs = fm + (QLs - QL) * sL[..., None], fp + (QRs - QR) * sR[..., None]
    sgn = np.sign(Bn)
    srL, srR = np.sqrt(rLs), np.sqrt(rRs)
    vyss = gdiv(vyRs * srR + vyLs * srL + sgn * (ByLs - ByRs), srL + srR)
    vzss = gdiv(vzRs * srR + vzLs * srL + sgn * (BzLs - BzRs), srL + srR)
    Byss = gdiv(ByLs * srR + ByRs * srL + sgn * (vyLs - vyRs) * np.sqrt(rRs * rLs), srL + srR)
    Bzss = gdiv(BzLs * srR + BzRs * srL + sgn * (vzLs - vzRs) * np.sqrt(rRs * rLs), srL + srR)

    def sstar(Qs, rs):
        S = np.zeros_like(Qs)
        S[..., 0] = rs
        S[..., a + 1] = sM
        S[..., o + 1] = vyss
        S[..., c + 1] = vzss
        S[..., a + 5] = np.copy(Qs[..., a + 5])
        S[..., o + 5] = Byss
        S[..., c + 5] = Bzss
        return S

    QLss, QRss = sstar(QLs, rLs), sstar(QRs, rRs)
    QLss[..., 4] = np.copy(QLs[..., 4] - np.sqrt(rLs) * np.sign(Bn) * (sum3(QLs[..., 1:4] * QLs[..., 5:8])
                                                                       - sum3(QLss[..., 1:4] * QLss[..., 5:8])))
    QRss[..., 4] = np.copy(QRs[..., 4] - np.sqrt(rRs) * np.sign(Bn) * (sum3(QRs[..., 1:4] * QRs[..., 5:8])
                                                                       - sum3(QRss[..., 1:4] * QRss[..., 5:8])))
    FLss, FRss = fm + (QLss - QLs) * sLs[..., None], fp + (QRss - QRs) * sRs[..., None]
    F = np.copy(fm)
    for m, src in [((sL <= 0) & (0 < sLs), FLs), ((sLs <= 0) & (0 < sM), FLss),
                   ((sM <= 0) & (0 < sRs), FRss), ((sRs <= 0) & (0 <= sR), FRs), (sR < 0, fp)]:
        F[m] = src[m]
    return F


def _fast_slow(rho, p, B, Bn_comp, gamma):
    cs = np.sqrt(gamma * gdiv(p, rho))
    ca = gdiv(np.sqrt(sum3(B * B)), np.sqrt(rho))
    cax = gdiv(Bn_comp, np.sqrt(rho))
    disc = np.sqrt(((cs ** 2 + ca ** 2) ** 2) - (4 * (cs ** 2) * (cax ** 2)))
    return cs, ca, cax, np.sqrt(.5 * (cs ** 2 + ca ** 2 + disc)), np.sqrt(.5 * (cs ** 2 + ca ** 2 - disc))


def right_eigenvectors(axis, g, gamma):
    """constructor.py:180-290: right eigenvectors of adiabatic MHD (Derigs et
    al.) evaluated on the states ``g`` (..., 8) read as (rho, v, p, B);
    returns (..., 8, 8), column k = wave k in the order fast-, Alfven-,
    slow-, entropy, divergence, slow+, Alfven+, fast+."""
    a, o, c = axis % 3, (axis + 1) % 3, (axis + 2) % 3
    rho, vel, p, B = g[..., 0], g[..., 1:4], g[..., 4], g[..., 5:8]
    vx, vy, vz = vel[..., a], vel[..., o], vel[..., c]
    Bx, By, Bz = B[..., a], B[..., o], B[..., c]
    cs, _, _, cf, csl = _fast_slow(rho, p, B, Bx, gamma)
    S = np.sign(Bx)
    S[S == 0] = 1
    af = np.sqrt(gdiv(cs ** 2 - csl ** 2, cf ** 2 - csl ** 2))
    als = np.sqrt(gdiv(cf ** 2 - cs ** 2, cf ** 2 - csl ** 2))
    bperp = np.sqrt(gdiv(By ** 2 + Bz ** 2, rho))
    b2 = gdiv(By, np.sqrt(By ** 2 + Bz ** 2))
    b3 = gdiv(Bz, np.sqrt(By ** 2 + Bz ** 2))
    v2 = norm2(vel)
    tr = vy * b2 + vz * b3
    # psi_{+-} (constructor.py:206-233): each a left-to-right sum of five terms
    base_s = (.5 * als * rho * v2 - cs * af * rho * bperp) + (als * rho * cs ** 2) / (gamma - 1)
    ts4, ts5 = als * csl * rho * vx, af * cf * rho * S * tr
    base_f = (.5 * af * rho * v2 + cs * als * rho * bperp) + (af * rho * cs ** 2) / (gamma - 1)
    tf4, tf5 = af * cf * rho * vx, als * csl * rho * S * tr
    r15, sr = rho ** 1.5, np.sqrt(rho)
    R = np.zeros(g.shape + (8,))
    # fast -/+ (columns 0 and 7, constructor.py:237-244, 275-282)
    for k, plus in ((0, False), (7, True)):
        R[..., 0, k] = rho * af
        R[..., o + 5, k] = als * cs * b2 * sr
        R[..., c + 5, k] = als * cs * b3 * sr
        if plus:
            R[..., a + 1, k] = rho * af * (vx + cf)
            R[..., o + 1, k] = rho * (af * vy - als * csl * b2 * S)
            R[..., c + 1, k] = rho * (af * vz - als * csl * b3 * S)
            R[..., 4, k] = (base_f + tf4) - tf5
        else:
            R[..., a + 1, k] = rho * af * (vx - cf)
            R[..., o + 1, k] = rho * (af * vy + als * csl * b2 * S)
            R[..., c + 1, k] = rho * (af * vz + als * csl * b3 * S)
            R[..., 4, k] = (base_f - tf4) + tf5
    # slow -/+ (columns 2 and 5, constructor.py:251-258, 267-274)
    for k, plus in ((2, False), (5, True)):
        R[..., 0, k] = rho * als
        R[..., o + 5, k] = -af * cs * b2 * sr
        R[..., c + 5, k] = -af * cs * b3 * sr
        if plus:
            R[..., a + 1, k] = rho * als * (vx + csl)
            R[..., o + 1, k] = rho * (als * vy + af * cf * b2 * S)
            R[..., c + 1, k] = rho * (als * vz + af * cf * b3 * S)
            R[..., 4, k] = (base_s + ts4) + ts5
        else:
            R[..., a + 1, k] = rho * als * (vx - csl)
            R[..., o + 1, k] = rho * (als * vy - af * cf * b2 * S)
            R[..., c + 1, k] = rho * (als * vz - af * cf * b3 * S)
            R[..., 4, k] = (base_s - ts4) - ts5
    # Alfven -/+ (constructor.py:245-250, 267-272)
    R[..., o + 1, 1], R[..., c + 1, 1] = -b3 * r15, b2 * r15
    R[..., 4, 1] = (b2 * vz - b3 * vy) * r15
    R[..., o + 1, 6], R[..., c + 1, 6] = b3 * r15, -b2 * r15
    R[..., 4, 6] = (b3 * vy - b2 * vz) * r15
    for k in (1, 6):
        R[..., o + 5, k] = -rho * b3
        R[..., c + 5, k] = rho * b2
    # entropy and divergence waves (constructor.py:259-266)
    R[..., 0, 3] = 1
    R[..., a + 1, 3], R[..., o + 1, 3], R[..., c + 1, 3] = vx, vy, vz
    R[..., 4, 3] = .5 * v2
    R[..., 4, 4] = Bx
    R[..., a + 5, 4] = 1
    return R


def gauss_legendre_unit():
    """generic.py:246-248: 3-point Gauss-Legendre nodes and weights mapped to
    [0, 1] (numpy leggauss, then x/2 + 1/2 and w/2)."""
    r, w = np.polynomial.legendre.leggauss(3)
    return .5 * r + .5, w / 2


def flux_dots(ax, gamma, d):
    """solvers.py:236-284: Osher-Solomon flux with the Dumbser-Toro path
    integral.  The path psi(s) = q- + s (q+ - q-) runs through the
    conservative states and is handed to the eigenvector routine and the wave
    speeds as if primitive (solvers.py:244-252); the speeds use index 1 and
    B[..., 0] whatever the axis (251-256), the eigenvectors the axis.
    |A| = R |Lambda| pinv(R) (271) with numpy's SVD pseudo-inverse (singular
    values <= 1e-15 sigma_max dropped), summed with the quadrature weights."""
    qp, qm = d["qFs"]
    fp, fm = d["fluxFs"]
    ax %= 3
    roots, weights = gauss_legendre_unit()
    P = np.repeat(qp[None], len(roots), axis=0)
    M = np.repeat(qm[None], len(roots), axis=0)
    psi = M + (roots * (P - M).T).T
    R = right_eigenvectors(ax, psi, gamma)
    rho, vx, p, B = psi[..., 0], psi[..., 1], psi[..., 4], psi[..., 5:8]
    _, _, cax, cf, csl = _fast_slow(rho, p, B, B[..., 0], gamma)
    lam = np.zeros_like(R)
    for k, v in enumerate((vx - cf, vx - cax, vx - csl, vx, vx, vx + csl, vx + cax, vx + cf)):
        lam[..., k, k] = v
    absA = R @ np.abs(lam) @ np.linalg.pinv(R)
    J = np.sum((weights * absA.T).T, axis=0)
    dq_p = (J @ qp[..., None])[..., 0]
    dq_m = (J @ qm[..., None])[..., 0]
    return .5 * (fp + fm) - .5 * (dq_p - dq_m)


def es_right_eigenvectors(axis, g, gamma):
    """constructor.py:294-415: the columns of right_eigenvectors (the same
    formulas), times the square root of Barth's diagonal scaling (404-414) as
    a matrix product, so a non-finite entry spreads exactly as there."""
    a, o, c = axis % 3, (axis + 1) % 3, (axis + 2) % 3
    R = right_eigenvectors(axis, g, gamma)
    rho, p = g[..., 0], g[..., 4]
    D = np.zeros_like(R)
    D[..., 0, 0] = 1 / (2 * gamma * rho)
    D[..., a + 1, a + 1] = gdiv(p, 2 * rho ** 2)
    D[..., o + 1, o + 1] = 1 / (2 * gamma * rho)
    D[..., c + 1, c + 1] = (rho * (gamma - 1)) / gamma
    D[..., 4, 4] = gdiv(p, rho)
    D[..., a + 5, a + 5] = 1 / (2 * gamma * rho)
    D[..., o + 5, o + 5] = gdiv(p, 2 * rho ** 2)
    D[..., c + 5, c + 5] = 1 / (2 * gamma * rho)
    return R @ np.sqrt(D)


def flux_es(ax, gamma, d):
    """solvers.py:288-391 entropy-stable flux (Winters & Gassner EC flux plus
    Derigs et al. hybrid Roe/LLF dissipation), as written: its "arithmetic
    mean" is the half-difference .5 (L - R) (306-307), the velocity and field
    components are taken at indices 0, 1, 2 whatever the axis (320-321), the
    eigenvectors at the averaged state with the axis.  The reference runs it
    in 1D only (its transpose(0, 2, 1) at 387 needs (N, 8, 8) stacks)."""
    wp, wm = d["wFs"]
    a, o, c = ax % 3, (ax + 1) % 3, (ax + 2) % 3

    def am(t):
        return .5 * (t[0] - t[1])

    def lon(t):
        return gdiv(t[0] - t[1], glog(t[0]) - glog(t[1]))

    rhoL, vL, PL, BL = wm[..., 0], wm[..., 1:4], wm[..., 4], wm[..., 5:8]
    rhoR, vR, PR, BR = wp[..., 0], wp[..., 1:4], wp[..., 4], wp[..., 5:8]
    z1 = np.array([np.sqrt(gdiv(rhoL, PL)), np.sqrt(gdiv(rhoR, PR))])
    z5 = np.array([np.sqrt(rhoL * PL), np.sqrt(rhoR * PR)])
    vx, vy, vz = (np.array([vL[..., i], vR[..., i]]) for i in range(3))
    Bx, By, Bz = (np.array([BL[..., i], BR[..., i]]) for i in range(3))
    rho_h = am(z1) * lon(z5)
    P1 = gdiv(am(z5), am(z1))
    P2 = ((gamma + 1) / (2 * gamma)) * (gdiv(lon(z5), lon(z1))) + ((gamma - 1) / (2 * gamma)) * (gdiv(am(z5), am(z1)))
    u1, v1, w1 = (gdiv(am(x * z1), am(z1)) for x in (vx, vy, vz))
    u2, v2, w2 = (gdiv(am(x * z1 ** 2), am(z1 ** 2)) for x in (vx, vy, vz))
    B1h, B1d, B2h, B2d, B3h, B3d = am(Bx), am(Bx ** 2), am(By), am(By ** 2), am(Bz), am(Bz ** 2)
    B1B2, B1B3 = am(Bx * By), am(Bx * Bz)
    ec = np.zeros_like(wp)
    ec[..., 0] = rho_h * u1
    ec[..., a + 1] = P1 + rho_h * u1 ** 2 + .5 * (B1d + B2d + B3d) - B1d
    ec[..., o + 1] = rho_h * u1 * v1 - B1B2
    ec[..., c + 1] = rho_h * u1 * w1 - B1B3
    ec[..., 4] = ((gamma / (gamma - 1)) * u1 * P2 + .5 * rho_h * u1 * (u1 ** 2 + v1 ** 2 + w1 ** 2)
                  + u2 * (B2h ** 2 + B3h ** 2) - B1h * (v2 * B2h + w2 * B3h))
    ec[..., o + 5] = u2 * B2h - v2 * B1h
    ec[..., c + 5] = u2 * B3h - w2 * B1h
    hat = np.stack([rho_h, u1, v1, w1, P1, B1h, B2h, B3h], axis=-1)
    R = es_right_eigenvectors(ax, hat, gamma)
    _, _, cax, cf, csl = _fast_slow(rho_h, P1, hat[..., 5:8], B1h, gamma)
    waves = (u1 + cf, u1 + cax, u1 + csl, u1, u1, u1 - csl, u1 - cax, u1 - cf)
    roe = np.zeros_like(R)
    for k, x in enumerate(waves):
        roe[..., k, k] = x
    roe = np.abs(roe)
    mx = np.maximum.reduce([np.abs(x) for x in waves[:4] + waves[5:]])
    lff = np.zeros_like(R)
    for k in range(8):
        lff[..., k, k] = mx
    e = np.zeros_like(wp)
    e[..., 0] = (((gamma - np.log(PL * rhoL ** -gamma)) / (gamma - 1) - gdiv(.5 * rhoL * norm2(vL), PL))
                 - ((gamma - np.log(PR * rhoR ** -gamma)) / (gamma - 1) - gdiv(.5 * rhoR * norm2(vR), PR)))
    e[..., 4] = gdiv(rhoR, PR) - gdiv(rhoL, PL)
    e[..., 1:4] = gdiv(vL * rhoL[..., None], PL[..., None]) - gdiv(vR * rhoR[..., None], PR[..., None])
    e[..., 5:8] = gdiv(BL * rhoL[..., None], PL[..., None]) - gdiv(BR * rhoR[..., None], PR[..., None])
    e *= -1
    eps = np.sqrt(np.abs(gdiv(PR - PL, PR + PL)))
    lam = (1 - eps)[..., None, None] * roe + eps[..., None, None] * lff
    absA = R @ lam @ R.transpose(0, 2, 1)
    diss = (absA @ e[..., None])[..., 0]
    return ec + .5 * diss


def riemann(s, ax, lmax, chars, d):
    """solvers.py:13-31 dispatch."""
    if s.solver_category == "hll":
        if s.solver.endswith("d"):
            return flux_hlld(ax, s.gamma, s.boundary, d)
        if s.compat == "astrea":
            return flux_hllc(ax, s.gamma, d)
        return flux_hllc(ax, s.gamma, d, ax + 1, default_minus=True)
    if s.solver_category == "complete":
        if s.solver.startswith("o"):
            return flux_dots(ax, s.gamma, d)
        return flux_es(ax, s.gamma, d)
    if s.solver.endswith("w"):
        if chars is not None:
            return flux_lw(chars, d)
        return flux_lw_speed(lw_sound_speed(d["jac_state"], s.gamma), lmax, d)
    return flux_llf(lmax, d)


# ----------------------------------------------------------------------------
# MHD constrained transport (num_methods/mag_field.py)
# ----------------------------------------------------------------------------
def transverse_reconstruct(wF, s, method="ppm"):
    """mag_field.py:11-121: reconstruct a sweep's face values along the other
    axis (ortho_axis of generic.py:264 = (1,0,2) in 2D) to the corners."""
    bc = s.boundary
    o = np.copy(wF.transpose(1, 0, 2))
    p2 = pad(o, bc, 2)
    p1 = np.copy(p2[1:-1])
    if method == "ppm":
        wU = C7_12 * (o + p1[2:]) - C1_12 * (p1[:-2] + p2[4:])
        wUp = pad(wU, bc, 2)
        wD, wU = mc_extrapolant_limiter(o, bc, np.copy(wUp[1:-3]), np.copy(wUp[2:-2]))
        return wD, wU
    eps = 1e-6
    z = np.copy(p1[1:-1])
    m1, m2, q1, q2 = p1[:-2], p2[:-4], p1[2:], p2[4:]
    g0, g1, g2 = 1 / 10, 3 / 5, 3 / 10
    b0 = 13 / 12 * (m2 - 2 * m1 + z) ** 2 + 1 / 4 * (m2 - 4 * m1 + 3 * z) ** 2
    b1 = 13 / 12 * (m1 - 2 * z + q1) ** 2 + 1 / 4 * (m1 - q1) ** 2
    b2 = 13 / 12 * (z - 2 * q1 + q2) ** 2 + 1 / 4 * (3 * z - 4 * q1 + q2) ** 2
    A = [lambda g, b=b: g / (b + eps) ** 2 for b in (b0, b1, b2)]
    sD = A[0](g2) + A[1](g1) + A[2](g0)
    wD = ((A[0](g2) / sD) * (1 / 3 * z + 5 / 6 * m1 - 1 / 6 * m2)
          + (A[1](g1) / sD) * (-1 / 6 * q1 + 5 / 6 * z + 1 / 3 * m1)
          + (A[2](g0) / sD) * (1 / 3 * q2 - 7 / 6 * q1 + 11 / 6 * z))
    sU = A[0](g0) + A[1](g1) + A[2](g2)
    wU = ((A[0](g0) / sU) * (1 / 3 * m2 - 7 / 6 * m1 + 11 / 6 * z)
          + (A[1](g1) / sU) * (-1 / 6 * m1 + 5 / 6 * z + 1 / 3 * q1)
          + (A[2](g2) / sU) * (1 / 3 * z + 5 / 6 * q1 - 1 / 6 * q2))
    return wD, wU


def corner_emf(wts, s):
    """mag_field.py:125-187.  ``wts`` maps physical axis -> (wD, wU) from the
    sweep of that axis.  The (north, south) / (east, west) roles follow the
    reference's dict iteration order, which flips with the step parity."""
    bc, gamma = s.boundary, s.gamma
    order = s.order()
    swapped = {k: perm(v, 2) for k, v in zip(order, order[::-1])}

    def speeds(wD, wU, ax):
        plus, minus = pad(wD, bc)[1:], pad(wU, bc)[:-1]
        avg = roe_average(plus, minus)[1:]
        if s.solver_category == "hll":
            rho, v, p, B = avg[..., 0], avg[..., 1:4], avg[..., 4], avg[..., 5:8]
            vx = v[..., ax % 3]
            _, _, _, cf, _ = _fast_slow(rho, p, B, B[..., ax % 3], gamma)
            return np.maximum(np.zeros_like(vx), vx + cf), -np.minimum(np.zeros_like(vx), vx - cf)
        ch = np.linalg.eigvals(jacobian(avg, gamma, ax % 3))
        return np.max(ch, axis=-1), -np.min(ch, axis=-1)

    alphas, comps = [], []
    for ax in order:
        axes = swapped[ax]
        wD, wU = wts[axes[0]]
        align = perm(ax, 2)
        ap, am = speeds(wD, wU, ax)
        alphas.append([ap.transpose(align[:-1]), am.transpose(align[:-1])])
        comps.append([wD.transpose(align), wU.transpose(align)])
    [north, south], [east, west] = comps
    NE = .5 * (west[..., 2] + south[..., 2]) * south[..., 5] - .5 * (west[..., 1] + south[..., 1]) * west[..., 6]
    NW = .5 * (east[..., 2] + south[..., 2]) * south[..., 5] - .5 * (east[..., 1] + south[..., 1]) * east[..., 6]
    SE = .5 * (west[..., 2] + north[..., 2]) * north[..., 5] - .5 * (west[..., 1] + north[..., 1]) * west[..., 6]
    SW = .5 * (east[..., 2] + north[..., 2]) * north[..., 5] - .5 * (east[..., 1] + north[..., 1]) * east[..., 6]
    [apy, amy], [apx, amx] = alphas
    return (gdiv(apx * apy * SW + amx * apy * SE + apx * amy * NW + amx * amy * NE, (apx + amx) * (apy + amy))
            - gdiv(apy * amy, apy + amy) * (north[..., 5] - south[..., 5])
            + gdiv(apx * amx, apx + amx) * (east[..., 6] - west[..., 6]))


def inverse_reconstruct(grid, s):
    """mag_field.py:191-210: face-averaged B back to cell averages."""
    bc = s.boundary
    new = np.copy(grid)
    for ax in s.order():
        axes = perm(ax, 2)
        rev = np.argsort(axes)
        fc = ho_convert("avg", grid.transpose(axes), bc, [1])
        p2 = pad(fc, bc, 2)
        p1 = np.copy(p2[1:-1])
        cc = -1 / 16 * (p1[:-2] + p2[4:]) + 9 / 16 * (fc + p1[2:])
        ca = ho_convert("cntr", cc, bc, [0, 1])
        new[..., 5 + ax] = ca.transpose(rev)[..., 5 + ax]
    return new


# ----------------------------------------------------------------------------
# evolve_space / evolve_time / driver (num_methods/evolvers.py, astrea.py)
# ----------------------------------------------------------------------------
def evolve_space(grid, s):
    """evolvers.py:12-34 + solvers.py:10-65.  Returns {physical axis:
    {'flux': sweep-frame interface fluxes, 'eigmax': float, ...}}."""
    order = s.order()
    data = {}
    for ax in order:
        axes = perm(ax, s.dimension)
        d = reconstruct(grid.transpose(axes), s, ax)
        if s.magnetic_2d:
            # pcm.py:24-25 (wS), plm.py:77-78 (wR), ppm.py:40-41 (wF), weno.py:167-168 (wR);
            # every call site uses the default method 'ppm' (mag_field.py:11)
            if s.subgrid in ("pcm", "constant", "c", "polynomial"):
                src = d["wS"]
            elif s.subgrid in ("ppm", "parabolic", "p"):
                src = d["wT_src"]
            else:
                src = d["wLR"][1]
            d["wTs"] = transverse_reconstruct(src, s, "ppm")
        data[ax] = d
    fluxes = {}
    for enum_ax, ax in enumerate(order):
        if s.compat == "physical":
            enum_ax = ax      # the solver sees the physical sweep axis
        d = data[ax]
        lmax, chars = local_max_eig(d["jac_state"], s, ax)
        eigmax = np.max(np.maximum(lmax[:-1], lmax[1:]))
        Favg = riemann(s, enum_ax, lmax, chars, d)
        if s.dimension == 2:
            hd = {"wS": d["wS"], "jac_state": d["jac_state"]}
            for key in ("wFs", "qFs", "fluxFs"):
                a, b = d[key]
                hd[key] = (ho_convert("avg", a, s.boundary, [1]), ho_convert("avg", b, s.boundary, [1]))
            Fc = riemann(s, enum_ax, lmax, chars, hd)
            F = np.copy(Fc) - C24 * lap(np.copy(Favg), s.boundary, 1)
        else:
            F = Favg
            if s.av and s.subgrid in ("ppm", "parabolic", "p"):
                F = F + av_flux(grid.transpose(perm(ax, 1)), d["wS"], ax, s)
        fluxes[ax] = {"flux": F, "eigmax": eigmax}
    if s.magnetic_2d:
        # evolvers.py:26-32: each sweep keeps its own wF; the corner EMF is
        # stored untransposed for the axis-1 sweep and transposed for axis 0.
        e3U = corner_emf({ax: data[ax]["wTs"] for ax in order}, s)
        for ax in order:
            fluxes[ax]["face_avg"] = data[ax]["wF"]
        fluxes[1]["mag_corner"] = np.copy(e3U)
        fluxes[0]["mag_corner"] = np.copy(e3U.T)
    return fluxes


def compute_L(fluxes, s):
    """evolvers.py:41-60."""
    total = 0
    for ax in s.order():
        axes = perm(ax, s.dimension)
        total += (np.diff(fluxes[ax]["flux"], axis=0) / s.dx).transpose(np.argsort(axes))
    if s.magnetic_2d:
        order = s.order()
        swapped = {k: perm(v, 2) for k, v in zip(order, order[::-1])}
        for ax in order:
            align = swapped[ax][:-1]
            pe = pad(fluxes[ax]["mag_corner"], s.boundary)
            mag = (-1) ** ax * np.diff(pe[1:], axis=0) / s.dx
            total[..., 5 + ax] = mag.transpose(align)
    return -total


def evolve_time(grid, fluxes, dt, s):
    """evolvers.py:38-206 (all stage schemes)."""
    L = lambda f: compute_L(f, s)
    refine = (lambda g: inverse_reconstruct(g, s)) if s.magnetic_2d else (lambda g: g)
    space = lambda g: evolve_space(g, s)
    L0 = L(fluxes)
    grid = np.copy(grid)
    if s.magnetic_2d:
        order = s.order()
        swapped = {k: perm(v, 2) for k, v in zip(order, order[::-1])}
        for ax in order:
            grid[..., 5 + ax] = fluxes[ax]["face_avg"].transpose(np.argsort(perm(ax, 2)))[..., 5 + ax]
    ts = s.timestep.lower()
    if ts.startswith("ssprk"):
        code = ts.replace(',', '').replace('(', '').replace(')', '').replace('ssprk', '')
        reg, order = int(code[:-1]), int(code[-1])
        if order == 4 and reg == 10:
            k = np.copy(grid)
            f = fluxes
            for _ in range(5):
                k += refine(1 / 6 * dt * L(f))
                f = space(k)
            k5 = refine(3 / 5 * grid + 6 / 15 * k + 1 / 15 * dt * L(f))
            f = space(k5)
            kk = np.copy(k5)
            for _ in range(4):
                kk += refine(1 / 6 * dt * L(f))
                f = space(kk)
            return refine(-11 / 35 * grid + 5 / 7 * k5 + 3 / 5 * kk + 1 / 10 * dt * L(f))
        if order == 4:
            k1 = refine(grid + .39175222657189 * dt * L0)
            f1 = space(k1)
            k2 = refine(.444370493651235 * grid + .555629506348765 * k1 + .368410593050371 * dt * L(f1))
            f2 = space(k2)
            k3 = refine(.620101851488403 * grid + .379898148511597 * k2 + .251891774271694 * dt * L(f2))
            f3 = space(k3)
            k4 = refine(.178079954393132 * grid + .821920045606868 * k3 + .544974750228521 * dt * L(f3))
            f4 = space(k4)
            return refine(.517231671970585 * k2 + .096059710526147 * k3 + .06369246866629 * dt * L(f3)
                          + .386708617503269 * k4 + .226007483236906 * dt * L(f4))
        if order == 3 and reg == 5:
            k1 = refine(grid + .3772689151171 * dt * L0)
            f1 = space(k1)
            k2 = refine(k1 + .3772689151171 * dt * L(f1))
            f2 = space(k2)
            k3 = refine(.56656131914033 * grid + .43343868085967 * k2 + .16352294089771 * dt * L(f2))
            f3 = space(k3)
            k4 = refine(.09299483444413 * grid + .0000209036962 * k1 + .90698426185967 * k3
                        + .00071997378654 * dt * L0 + .34217696850008 * dt * L(f3))
            f4 = space(k4)
            return refine(.0073613226092 * grid + .20127980325145 * k1 + .00182955389682 * k2
                          + .78952932024253 * k4 + (dt * (.0027771981946 * L0 + .00001567934613 * L(f1)
                                                          + .29786487010104 * L(f4))))
        if order == 3 and reg == 4:
            k1 = refine(grid + .5 * dt * L0)
            f1 = space(k1)
            k2 = refine(k1 + .5 * dt * L(f1))
            f2 = space(k2)
            k3 = refine(1 / 6 * (4 * grid + 2 * k2 + dt * L(f2)))
            f3 = space(k3)
            return refine(k3 + .5 * dt * L(f3))
        if order == 3:
            k1 = refine(grid + dt * L0)
            f1 = space(k1)
            k2 = refine(.25 * (3 * grid + k1 + dt * L(f1)))
            f2 = space(k2)
            return refine(1 / 3 * (grid + 2 * k2 + 2 * dt * L(f2)))
        k1 = refine(grid + dt * L0)
        f1 = space(k1)
        return refine(.5 * (grid + k1 + dt * L(f1)))
    if ts.startswith("r"):
        k1 = refine(grid + .5 * dt * L0)
        f1 = space(k1)
        k2 = refine(grid + .5 * dt * L(f1))
        f2 = space(k2)
        k3 = refine(grid + dt * L(f2))
        f3 = space(k3)
        return refine(grid + 1 / 6 * (dt * (L0 + 2 * L(f1) + 2 * L(f2) + L(f3))))
    return refine(grid + dt * L0)


@dataclass
class RunResult:
    grid: np.ndarray
    t: float
    steps: int
    dts: list = field(default_factory=list)
    clamped: list = field(default_factory=list)   # steps (1-based) whose dt was clamped to a checkpoint


def snapshot(grid, s):
    """astrea.py:47: convert_conservative of the grid, transposed by
    ortho_axis (generic.py:264: the identity in 1D, (1, 0, 2) in 2D)."""
    w = to_prim(grid, s)
    return w.transpose(1, 0, 2) if s.dimension == 2 else w


def run(grid, s, t_end, checkpoints=1, max_steps=None, snapshots=None):
    """astrea.py:31-85 time loop: dt from the stage-0 eigmax, clamp to the
    next checkpoint, exact-equality stop, sweep order flip.  With a list
    ``snapshots``, each pass of the loop appends (str(float(t)), snapshot)
    before the stop test, as astrea.py:46-50 writes one HDF5 dataset per pass."""
    chkpt = t_end / checkpoints
    t, idx, dts, n = 0., 1, [], 0
    clamped = []
    s = replace(s, parity=0)
    while t <= t_end:
        if snapshots is not None:
            snapshots.append((str(float(t)), snapshot(grid, s)))
        if t == t_end or (max_steps is not None and n >= max_steps):
            break
        fl = evolve_space(grid, s)
        dt = s.cfl * min([s.dx / fl[ax]["eigmax"] for ax in s.order()])
        if t + dt >= chkpt * idx:
            dt = chkpt * idx - t
            idx += 1
            clamped.append(n + 1)
        grid = evolve_time(grid, fl, dt, s)
        t += dt
        dts.append(dt)
        n += 1
        s = replace(s, parity=1 - s.parity)
    return RunResult(grid, t, n, dts, clamped)


def initial_average(w_point, s):
    """constructor.py:105-109: pointwise primitive IC -> conservative, then
    +h^2/24 Laplacian for every scheme."""
    q = to_cons(w_point, replace(s, parity=0), "cell")
    return ho_convert("cntr", q, s.boundary, range(s.dimension))


# ----------------------------------------------------------------------------
# Diagnostics of snapshots (functions/analytic.py); a snapshot is the
# primitive array the reference stores each step (astrea.py:47)
# ----------------------------------------------------------------------------
def _divide(a, b):
    """fv.divide (fv.py:19-20): a / b where b != 0, else 0."""
    return np.divide(a, b, out=np.zeros_like(a), where=b != 0)


def snapshot_energies(w, gamma):
    """Specific total and internal energy of a primitive snapshot
    (analytic.py:33-34, 54-55; fv.convert_variable('pressure'), fv.py:51)."""
    e = w[..., 4] / (gamma - 1) + .5 * (w[..., 0] * norm2(w[..., 1:4]) + norm2(w[..., 5:8]))
    return _divide(e, w[..., 0]), _divide(w[..., 4], w[..., 0] * (gamma - 1))


def solution_error(w_num, w_theo, gamma, p):
    """analytic.py:24-44 (10 components: the 8 primitives, E_tot, E_int)."""
    dims = w_num.ndim - 1
    fac = 1 / (len(w_num) ** dims)
    et_n, ei_n = snapshot_energies(w_num, gamma)
    et_t, ei_t = snapshot_energies(w_theo, gamma)
    a = np.concatenate((w_num, et_n[..., None], ei_n[..., None]), axis=-1)
    b = np.concatenate((w_theo, et_t[..., None], ei_t[..., None]), axis=-1)
    ax = tuple(range(dims))
    if p > 10:
        return np.max(np.abs(a - b), axis=ax)
    if p <= 0:
        return fac * np.sum(np.abs(a - b), axis=ax)
    return (fac * np.sum(np.abs(a - b) ** p, axis=ax)) ** (1 / p)


def total_variation(w, gamma):
    """analytic.py:48-62 for one snapshot: the successive differences along
    every axis, summed in absolute value (8 primitives, E_tot, E_int)."""
    dims = w.ndim - 1
    et, ei = snapshot_energies(w, gamma)
    g = np.copy(w)
    for i in range(dims):
        g, et, ei = np.diff(g, axis=i), np.diff(et, axis=i), np.diff(ei, axis=i)
    tv = np.sum(np.abs(g), axis=tuple(range(dims)))
    return np.append(np.append(tv, np.sum(np.abs(et))), np.sum(np.abs(ei)))


def conservation(w, s, box_width=1.0):
    """analytic.py:66-75 for one snapshot: convert_primitive (4th-order for
    PPM/WENO, along the snapshot's own axes), summed, times box_width^dim."""
    q = to_cons(w, s)
    return np.sum(q, axis=tuple(range(w.ndim - 1))) * box_width ** (w.ndim - 1)

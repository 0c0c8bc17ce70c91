"""Exact solution of the 1D Riemann problem of the Euler equations for an
ideal gas (Toro, "Riemann Solvers and Numerical Methods for Fluid
Dynamics", 3rd ed., chapter 4: the pressure function of section 4.2, the
Newton iteration of 4.3 and the sampling of 4.5).  Test helper: the
reference solves only Sod's problem exactly (functions/analytic.py:98-153),
which assumes a left state at rest."""
import numpy as np


def _pressure_function(p, rho, pk, ck, gamma):
    """f_K(p) and f_K'(p) (Toro eq. 4.6-4.7, 4.37)."""
    if p > pk:                                   # shock
        a = 2 / ((gamma + 1) * rho)
        b = (gamma - 1) / (gamma + 1) * pk
        q = np.sqrt(a / (b + p))
        return (p - pk) * q, q * (1 - .5 * (p - pk) / (b + p))
    r = p / pk                                   # rarefaction
    return (2 * ck / (gamma - 1) * (r ** ((gamma - 1) / (2 * gamma)) - 1),
            r ** (-(gamma + 1) / (2 * gamma)) / (rho * ck))


def star_state(left, right, gamma):
    """(p*, u*) for primitive states (rho, u, p) left and right."""
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)
    p = max(1e-8, .5 * (pl + pr) - .125 * (ur - ul) * (rl + rr) * (cl + cr))   # PVRS start (eq. 4.47)
    for _ in range(100):
        fl, dl = _pressure_function(p, rl, pl, cl, gamma)
        fr, dr = _pressure_function(p, rr, pr, cr, gamma)
        p_new = max(1e-12, p - (fl + fr + ur - ul) / (dl + dr))
        if abs(p_new - p) <= 1e-14 * .5 * (p_new + p):
            p = p_new
            break
        p = p_new
    fl, _ = _pressure_function(p, rl, pl, cl, gamma)
    fr, _ = _pressure_function(p, rr, pr, cr, gamma)
    return p, .5 * (ul + ur) + .5 * (fr - fl)


def sample(left, right, gamma, xi):
    """Primitive (rho, u, p) at similarity coordinates xi = (x - x0) / t."""
    ps, us = star_state(left, right, gamma)
    g1, g2 = (gamma - 1) / (2 * gamma), (gamma + 1) / (2 * gamma)
    g6 = (gamma - 1) / (gamma + 1)
    out = np.zeros((len(xi), 3))
    for i, s in enumerate(np.asarray(xi, dtype=float)):
        side, sgn = (left, 1.) if s <= us else (right, -1.)
        rk, uk, pk = side
        ck = np.sqrt(gamma * pk / rk)
        if ps > pk:                                            # shock on this side
            sk = uk - sgn * ck * np.sqrt(g2 * ps / pk + g1)
            if (s - sk) * sgn <= 0:
                out[i] = side
            else:
                out[i] = (rk * (ps / pk + g6) / (g6 * ps / pk + 1), us, ps)
        else:                                                  # rarefaction on this side
            head = uk - sgn * ck
            tail = us - sgn * ck * (ps / pk) ** g1
            if (s - head) * sgn <= 0:
                out[i] = side
            elif (s - tail) * sgn >= 0:
                out[i] = (rk * (ps / pk) ** (1 / gamma), us, ps)
            else:                                              # inside the fan
                u = 2 / (gamma + 1) * (sgn * ck + (gamma - 1) / 2 * uk + s)
                c = 2 / (gamma + 1) * (ck + sgn * (gamma - 1) / 2 * (uk - s))
                out[i] = (rk * (c / ck) ** (2 / (gamma - 1)), u, pk * (c / ck) ** (2 * gamma / (gamma - 1)))
    return out

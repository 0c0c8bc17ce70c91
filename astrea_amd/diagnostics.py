"""Diagnostics of device snapshots (functions/analytic.py:24-153).

A snapshot is the primitive array the reference stores every step
(``Plan.snapshot``: convert_conservative, transposed by ortho_axis); the
diagnostics here read it where it lies, with FP64 torch reductions on the
device, and return small host arrays.  They follow analytic.py's formulas;
summation order differs from numpy's pairwise sums, so values agree with the
reference's to rounding, not bit for bit.
"""
import math

import torch

from . import ic


def _divide(a, b):
    """fv.divide (fv.py:19-20): a / b where b != 0, else 0."""
    return torch.where(b != 0, a / torch.where(b != 0, b, torch.ones_like(b)), torch.zeros_like(a))


def energies(w, gamma):
    """Specific total and internal energy (analytic.py:33-34, 54-55)."""
    e = w[..., 4] / (gamma - 1) + .5 * (w[..., 0] * ic._norm2(w[..., 1:4]) + ic._norm2(w[..., 5:8]))
    return _divide(e, w[..., 0]), _divide(w[..., 4], w[..., 0] * (gamma - 1))


def solution_error(w_num, w_theo, gamma, p):
    """L^p error of a snapshot against a theoretical primitive array
    (analytic.py:24-44): 10 values, the 8 primitives, E_tot and E_int;
    p > 10 is the maximum norm, p <= 0 the normalised L1 sum."""
    w_theo = torch.as_tensor(w_theo, dtype=torch.float64, device=w_num.device)
    dims = w_num.dim() - 1
    fac = 1 / (w_num.shape[0] ** dims)
    et_n, ei_n = energies(w_num, gamma)
    et_t, ei_t = energies(w_theo, gamma)
    d = (torch.cat((w_num, et_n[..., None], ei_n[..., None]), -1) -
         torch.cat((w_theo, et_t[..., None], ei_t[..., None]), -1)).abs()
    ax = tuple(range(dims))
    if p > 10:
        out = d.amax(dim=ax)
    elif p <= 0:
        out = fac * d.sum(dim=ax)
    else:
        out = (fac * (d ** p).sum(dim=ax)) ** (1 / p)
    return out.cpu().numpy()


def total_variation(w, gamma):
    """Total variation of a snapshot (analytic.py:48-62): successive
    differences along every axis, summed in absolute value; 10 values."""
    dims = w.dim() - 1
    et, ei = energies(w, gamma)
    g = w
    for i in range(dims):
        g, et, ei = torch.diff(g, dim=i), torch.diff(et, dim=i), torch.diff(ei, dim=i)
    tv = g.abs().sum(dim=tuple(range(dims)))
    return torch.cat((tv, et.abs().sum()[None], ei.abs().sum()[None])).cpu().numpy()


def conservation(w, gamma, boundary, high_order, box_width=1.0):
    """Totals of the conservative variables of a snapshot (analytic.py:66-75):
    convert_primitive along the snapshot's own axes (4th-order for PPM/WENO),
    summed, times box_width^dim; 8 values."""
    dims = w.dim() - 1
    if high_order:
        ww, q = w.clone(), torch.zeros_like(w)
        pw = ic.prim_to_cons(w, gamma)
        for ax in range(dims):
            ww -= ic.K24 * ic._lap(w, ax, boundary)
            q += ic.K24 * ic._lap(pw, ax, boundary)
        q = ic.prim_to_cons(ww, gamma) + q
    else:
        q = ic.prim_to_cons(w, gamma)
    return (q.sum(dim=tuple(range(dims))) * box_width ** dims).cpu().numpy()


def div_b(plan, i):
    """max |div B| of a 2D MHD problem by cell-centred central differences
    (the reference has no divergence diagnostic), and the same divided by
    max|B| / dx; periodic neighbours."""
    n0, n1 = plan.shape(i)
    o = plan.offsets[i]
    bx = plan.state[plan.varidx.index(5), o:o + n0 * n1].reshape(n0, n1)
    by = plan.state[plan.varidx.index(6), o:o + n0 * n1].reshape(n0, n1)
    dx = plan.problems[i].dx
    div = (torch.roll(bx, -1, 0) - torch.roll(bx, 1, 0)) / (2 * dx) + \
          (torch.roll(by, -1, 1) - torch.roll(by, 1, 1)) / (2 * dx)
    m = float(div.abs().max())
    bmax = float(torch.sqrt(bx * bx + by * by).max())
    return m, m * dx / bmax if bmax > 0 else math.inf


def _round_off(value):
    """analytic.py:10-14 (rounds half up, away from the integer part)."""
    if value % int(value) >= .5:
        return int(value) + 1
    return int(value)


def sod_exact(w_grid, t, gamma, shock_pos=.5, start=0., end=1.):
    """Exact Sod solution on the cells of a 1D primitive array
    (analytic.py:98-153): the end cells give the initial left (5) and right (1)
    states, the post-shock pressure solves the reference's root equation
    (scipy fsolve from (P5 - P1)/2), and the regions are laid out by the
    reference's rounded cell boundaries.  Host-side (numpy/scipy)."""
    import numpy as np
    from scipy import optimize
    grid = np.asarray(w_grid, dtype=np.float64)
    arr = np.zeros_like(grid)
    rho5, vx5, vy5, vz5, P5 = grid[0][:5]
    rho1, vx1, vy1, vz1, P1 = grid[-1][:5]
    cs5, cs1 = np.sqrt(gamma * P5 / rho5), np.sqrt(gamma * P1 / rho1)
    mu, beta = (gamma - 1) / (gamma + 1), 2 / (gamma - 1)
    f = lambda x: (((x / P1) - 1) * np.sqrt((1 - mu) / (gamma * (mu + (x / P1))))) - \
        (beta * (cs5 / cs1) * (1 - ((x / P5) ** (1 / (gamma * beta)))))
    P2 = optimize.fsolve(f, (P5 - P1) / 2)[0]
    rho2, rho3 = rho1 * ((P2 + (mu * P1)) / (P1 + (mu * P2))), rho5 * (P2 / P5) ** (1 / gamma)
    vx2 = (beta * cs5) * (1 - (P2 / P5) ** (1 / (gamma * beta)))
    v_t = cs5 - (vx2 / (1 - mu))
    v_s = vx2 / (1 - (rho1 / rho2))
    n = len(grid)
    b54 = _round_off(((shock_pos - (cs5 * t) - start) / (end - start)) * n)
    b43 = _round_off(((shock_pos - (v_t * t) - start) / (end - start)) * n)
    b32 = _round_off(((shock_pos + (vx2 * t) - start) / (end - start)) * n)
    b21 = _round_off(((shock_pos + (v_s * t) - start) / (end - start)) * n)
    rare = _round_off(((cs5 * t - v_t * t) / (end - start)) * n)
    if rare - (b43 - b54) < 0:
        rare += 1
    elif rare - (b43 - b54) > 0:
        rare -= 1
    x = np.linspace(shock_pos - (cs5 * t), shock_pos - (v_t * t), rare) - shock_pos
    arr[:b54] = grid[0]
    arr[b21:] = grid[-1]
    arr[b43:b21, 1] = vx2
    arr[b43:b21, 4] = P2
    arr[b43:b32, 0] = rho3
    arr[b32:b21, 0] = rho2
    arr[b54:b43, 0] = rho5 * ((1 - mu) - mu * x / (cs5 * t)) ** beta
    arr[b54:b43, 4] = P5 * ((1 - mu) - mu * x / (cs5 * t)) ** (gamma * beta)
    arr[b54:b43, 1] = (1 - mu) * (cs5 + (x / t))
    return arr

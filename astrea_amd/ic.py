"""Initial conditions.

The reference builds a pointwise primitive state, converts it to
conservative variables and adds the h^2/24 Laplacian for every scheme
(functions/constructor.py:105-109); the conversion is the 4th-order one for
PPM/WENO and the pointwise one otherwise (generic.py:250-255).  Here the same
pipeline runs in FP64 torch on the host, each elementwise operation rounding
once exactly as numpy does (pinned against the oracle's restatement), and
``to_device`` copies the result.  torch's GPU elementwise kernels differ from
numpy in the last bit on some cells, so the averaging does not run there;
piecewise-constant problems are instead built on the device by
ah_plan_init_ic, bit for bit this pipeline.  Generators for the benchmark
problems draw their random numbers on the host with numpy's PCG64 (seed
stated by the caller) so every run is reproducible.
"""
import math

import numpy as np
import torch

K24 = 1 / 24


def _lap(x, axis, bc):
    n = x.shape[axis]
    idx = torch.arange(n, device=x.device)
    if bc == "wrap":
        up, lo = (idx + 1) % n, (idx - 1) % n
    else:
        up, lo = torch.clamp(idx + 1, max=n - 1), torch.clamp(idx - 1, min=0)
    xu, xl = x.index_select(axis, up), x.index_select(axis, lo)
    return (xu - x) - (x - xl)


def _norm2(v):
    n = torch.sqrt((v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2])
    return n * n


def prim_to_cons(w, gamma):
    q = w.clone()
    q[..., 4] = w[..., 4] / (gamma - 1) + .5 * (w[..., 0] * _norm2(w[..., 1:4]) + _norm2(w[..., 5:8]))
    q[..., 1:4] = w[..., 1:4] * w[..., 0][..., None]
    return q


def initial_average(w_point, gamma, bc, high_order):
    """Pointwise primitive (device tensor, (cells..., 8)) -> averaged conservative."""
    dims = w_point.dim() - 1
    if high_order:
        ww, q = w_point.clone(), torch.zeros_like(w_point)
        pw = prim_to_cons(w_point, gamma)
        for ax in range(dims):
            ww -= K24 * _lap(w_point, ax, bc)
            q += K24 * _lap(pw, ax, bc)
        q = prim_to_cons(ww, gamma) + q
    else:
        q = prim_to_cons(w_point, gamma)
    out = q.clone()
    for ax in range(dims):
        out += K24 * _lap(q, ax, bc)
    return out


def cell_centres(n, start=0., end=1.):
    """constructor.py:13-16 (linspace over the ghost-extended edges)."""
    dx = abs(end - start) / n
    return np.linspace(start - dx / 2, end + dx / 2, n + 2)[1:-1]


def sod(n):
    x = cell_centres(n)
    w = np.zeros((n, 8))
    w[:] = [.125, 0, 0, 0, .1, 0, 0, 0]
    w[x <= .5] = [1, 0, 0, 0, 1, 0, 0, 0]
    return w


def shear_layer(n, seed):
    """2D shear layer: rho=2, vx=+0.5 for |y-0.5|<0.25, else rho=1, vx=-0.5;
    p=2.5; vy = 0.01 sin(4 pi x); uniform noise of amplitude 1e-3 drawn for
    vx then vy, each an (n, n) array in C order over (i, j)."""
    rng = np.random.default_rng(seed)
    c = cell_centres(n)
    x, y = np.meshgrid(c, c, indexing="ij")
    w = np.zeros((n, n, 8))
    band = np.abs(y - .5) < .25
    w[..., 0] = np.where(band, 2., 1.)
    w[..., 1] = np.where(band, .5, -.5)
    w[..., 4] = 2.5
    w[..., 2] = .01 * np.sin(4 * np.pi * x)
    w[..., 1] += rng.uniform(-1e-3, 1e-3, (n, n))
    w[..., 2] += rng.uniform(-1e-3, 1e-3, (n, n))
    return w


def orszag_tang(n):
    c = cell_centres(n)
    x, y = np.meshgrid(c, c, indexing="ij")
    w = np.zeros((n, n, 8))
    w[..., 0] = 25 / (36 * np.pi)
    w[..., 4] = 5 / (12 * np.pi)
    w[..., 1] = -np.sin(2 * np.pi * y)
    w[..., 2] = np.sin(2 * np.pi * x)
    w[..., 5] = -np.sin(2 * np.pi * y) / math.sqrt(4 * np.pi)
    w[..., 6] = np.sin(4 * np.pi * x) / math.sqrt(4 * np.pi)
    return w


def riemann_1d(n, left, right, x0=.5):
    """1D Riemann problem; left/right = (rho, u, p)."""
    x = cell_centres(n)
    w = np.zeros((n, 8))
    w[:, 0], w[:, 1], w[:, 4] = right
    m = x <= x0
    w[m, 0], w[m, 1], w[m, 4] = left
    return w


def quadrants(n, states):
    """2D four-quadrant problem; states[k] = (rho, u, v, p) for quadrants
    k = 0 (x<=.5, y<.5), 1 (x>.5, y<.5), 2 (x<=.5, y>=.5), 3 (x>.5, y>=.5)
    (the Lax-Liu layout of constructor.py:57-60)."""
    c = cell_centres(n)
    x, y = np.meshgrid(c, c, indexing="ij")
    w = np.zeros((n, n, 8))
    masks = [(x <= .5) & (y < .5), (x > .5) & (y < .5), (x <= .5) & (y >= .5), (x > .5) & (y >= .5)]
    for m, (r, u, v, p) in zip(masks, states):
        w[m, 0], w[m, 1], w[m, 2], w[m, 4] = r, u, v, p
    return w


def ensemble_1d(nprob, seed, nmin=256, nmax=8192):
    """Seeded 1D Riemann ensemble: per problem, in this order, N log-uniform
    in [nmin, nmax], rho_L, rho_R log-uniform in [0.1, 10], p_L, p_R
    log-uniform in [0.01, 100], u_L, u_R uniform in [-2, 2], t_end uniform in
    [0.1, 0.25]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nprob):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        n -= n % 2
        rl, rr = np.exp(rng.uniform(math.log(.1), math.log(10), 2))
        pl, pr = np.exp(rng.uniform(math.log(.01), math.log(100), 2))
        ul, ur = rng.uniform(-2, 2, 2)
        t_end = rng.uniform(.1, .25)
        out.append(dict(n=n, left=(rl, ul, pl), right=(rr, ur, pr), t_end=t_end))
    return out


def quadrant_requests(nreq, seed, sizes=(128, 256, 512, 1024)):
    """Seeded four-quadrant requests: grid size uniform from ``sizes``,
    per quadrant rho log-uniform in [0.1, 2], p log-uniform in [0.05, 1.5],
    u and v uniform in [-1, 1]; PLM-minmod or PPM by coin flip."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nreq):
        n = int(sizes[rng.integers(len(sizes))])
        st = []
        for _q in range(4):
            r = math.exp(rng.uniform(math.log(.1), math.log(2)))
            p = math.exp(rng.uniform(math.log(.05), math.log(1.5)))
            u, v = rng.uniform(-1, 1, 2)
            st.append((r, u, v, p))
        out.append(dict(n=n, states=st, subgrid="plm" if rng.random() < .5 else "ppm"))
    return out


def initial_point(sv):
    """constructor.py:11-103: the pointwise primitive state of configuration
    sv.config (tables of ``configs.test_conditions`` in sv) on sv.cells cells
    per axis, in numpy FP64 with the reference's operations in its order."""
    config, N, gamma, dim = sv.config, sv.cells, sv.gamma, sv.dimension
    lo, hi, x0, prm = sv.start_pos, sv.end_pos, sv.shock_pos, sv.misc
    left = np.asarray(sv.initial_left)
    grid = np.zeros((N,) * dim + (len(sv.initial_right),))
    grid[:] = sv.initial_right
    h = abs(hi - lo) / N
    xs = np.linspace(lo - h / 2, hi + h / 2, N + 2)[1:-1]
    if dim == 2:
        x, y = np.meshgrid(xs, xs, indexing="ij")
        c = (hi + lo) / 2
        if config == "sedov" or "blast" in config:
            grid[((x - c) ** 2 + (y - c) ** 2) <= (x0 - c) ** 2] = left
        elif config.startswith("gauss"):
            r = np.sqrt((x - c) ** 2 + (y - c) ** 2)
            grid[..., 0] = prm["y_offset"] + prm["ampl"] * np.exp(-((r - c) ** 2) / prm["fwhm"])
        elif config in ("khi", "kelvin-helmholtz") or "kelvin" in config or "helmholtz" in config:
            grid[y <= x0] = left
            grid[..., 2] = prm["perturb_ampl"] * np.sin(prm["freq"] * np.pi * x / (hi - lo))
        elif config in ("ivc", "vortex", "isentropic vortex"):
            xc, yc = (np.min(x) + np.max(x)) / 2, (np.min(y) + np.max(y)) / 2
            r = np.sqrt((x - xc) ** 2 + (y - yc) ** 2)
            k = prm["freq"] * np.pi
            T = 1 - (((gamma - 1) * prm["vortex_str"] ** 2) / (2 * gamma * k ** 2)) * np.exp(1 - r ** 2)
            grid[..., 0] = T ** (1 / (gamma - 1))
            grid[..., 1] = (prm["vortex_str"] / k) * np.exp((1 - r ** 2) / 2)
            grid[..., 2] = (prm["vortex_str"] / k) * np.exp((1 - r ** 2) / 2)
            grid[..., 4] = T ** (gamma / (gamma - 1))
        elif "ll" in config or "lax-liu" in config:
            grid[x <= x0] = left
            grid[(x <= x0) & (y >= x0)] = prm["bottom_left"]
            grid[(x > x0) & (y >= x0)] = prm["bottom_right"]
        elif config in ("orszag-tang", "orszag", "tang", "ot"):
            grid[y <= x0] = left
            grid[..., 1] = -np.sin(2 * np.pi * y)
            grid[..., 2] = np.sin(2 * np.pi * x)
            grid[..., 5] = -prm["ampl"] * np.sin(2 * np.pi * y)
            grid[..., 6] = prm["ampl"] * np.sin(4 * np.pi * x)
        elif "rotor" in config:
            # constructor.py:79-81 writes the rotation into a fancy-indexed
            # copy, so the rotor keeps zero velocity
            grid[((x - c) ** 2 + (y - c) ** 2) <= (x0 - c) ** 2] = left
        else:
            grid[x < x0] = left
    else:
        x = xs
        grid[(np.abs(x) <= x0) if (config == "sedov" or config.startswith("sq")) else (x <= x0)] = left
        if "shu" in config or "osher" in config:
            xr = x[x > x0]
            grid[x > x0, 0] = prm["y_offset"] + prm["ampl"] * np.sin(prm["freq"] * np.pi * xr)
        elif config.startswith("sin"):
            grid[..., 0] = prm["y_offset"] + prm["ampl"] * np.sin(prm["freq"] * np.pi * x)
        elif config.startswith("gauss"):
            grid[..., 0] = prm["y_offset"] + prm["ampl"] * np.exp(-((x - prm["peak_pos"]) ** 2) / prm["fwhm"])
        elif config.startswith("lin"):
            if "mhd" in config:   # Alfven wave eigenvector (constructor.py:94-101)
                pert = np.array([0, 0, -.3333333333333333, .9428090415820634, 0, -.3333333333333333,
                                 .9428090415820634, 0]) * prm["ampl"]
            else:
                pert = np.array([1, -1, 1, 1, 1.5, 0, 0, 0]) * prm["ampl"]
            grid += pert * np.sin(prm["freq"] * np.pi * x)[..., None]
    return grid


def initialise(sv):
    """constructor.initialise(sv, convert=True) (constructor.py:11-109): the
    averaged conservative initial state of sv.config, as numpy FP64."""
    high = sv.subgrid.startswith("w") or sv.subgrid in ("ppm", "parabolic", "p")
    w = torch.from_numpy(initial_point(sv))
    return initial_average(w, sv.gamma, sv.boundary, high).numpy()


def to_device(w_point, gamma, bc, high_order, device="cuda:0"):
    """Averaged conservative initial state (host pipeline), on ``device``."""
    w = torch.as_tensor(w_point, dtype=torch.float64, device="cpu")
    return initial_average(w, gamma, bc, high_order).to(device)

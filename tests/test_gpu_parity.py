"""Parity of the HIP path (through the C ABI) with the oracle and with the
reference's golden trajectories.  Every case prints its name before touching
the GPU."""
import sys
from dataclasses import replace

import numpy as np
import pytest

from oracle import astrea_oracle as O
from tests._golden import load, meta, rel_l1

pytestmark = pytest.mark.gpu

HYDRO = sorted(k for k, v in meta()["cases"].items() if k != "rp1d_fail" and not v.get("magnetic_2d"))


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_run(grid0, m, precision="fp64", chunk=16):
    from astrea_amd.plan import Plan, infer_varmask
    dims = m["dimension"]
    mask = infer_varmask([grid0], dims, m["magnetic_2d"])
    prob = dict(n0=grid0.shape[0], n1=grid0.shape[1] if dims == 2 else 1, boundary=m["boundary"], dx=m["dx"],
                gamma=m["gamma"], cfl=m["cfl"], t_end=m["t_end"], checkpoints=m["checkpoints"],
                max_steps=m["max_steps"] if m["max_steps"] is not None else -1)
    plan = Plan(dims, m["subgrid"], m["solver"], m["timestep"], [prob], varmask=mask, precision=precision)
    plan.load([grid0])
    plan.reset()
    st = plan.run(chunk)
    out = plan.unpack(0)
    plan.close()
    return out, st[0]


def oracle_run(grid0, m, eig="closed"):
    s = O.Scheme(dimension=m["dimension"], subgrid=m["subgrid"], solver=m["solver"], timestep=m["timestep"],
                 boundary=m["boundary"], gamma=m["gamma"], cfl=m["cfl"], dx=m["dx"], eig=eig)
    return O.run(grid0, s, m["t_end"], m["checkpoints"], m["max_steps"])


@pytest.mark.parametrize("name", HYDRO)
def test_trajectory_parity(name):
    """Same step count as the reference; relative L1 per variable <= 1e-12
    against the oracle (closed-form eigenvalues, as on the device) and
    <= 1e-12 against the reference's own trajectory (WENO-7: 1e-10, its
    weights amplify the ulp-level LAPACK differences)."""
    m, g = meta()["cases"][name], load("steps.npz")
    grid0 = g[name + "_grid0"]
    _say(f"[gpu] trajectory {name} shape={grid0.shape} steps={m['steps']}")
    out, st = gpu_run(grid0, m)
    ref = oracle_run(grid0, m)
    assert not st["failed"]
    assert st["steps"] == m["steps"] == ref.steps
    e_or = np.max(rel_l1(out, ref.grid))
    e_gold = np.max(rel_l1(out, g[name + "_final"]))
    _say(f"[gpu] {name}: rel L1 vs oracle {e_or:.2e}, vs reference {e_gold:.2e}")
    assert e_or <= 1e-12
    assert e_gold <= (1e-10 if "weno7" in name else 1e-12)


SCHEME_KEYS = [f"hd_{sg}_{bc}" for sg in ("pcm", "plm", "ppm", "weno3", "weno5", "weno7") for bc in ("edge", "wrap")]


@pytest.mark.parametrize("key", SCHEME_KEYS)
@pytest.mark.parametrize("solver", ["llf", "hllc"])
def test_interface_flux_1d(key, solver):
    """ah_evolve_space interface fluxes and eigmax against the reference's
    solver outputs (schemes_1d.npz) and the oracle."""
    from astrea_amd.plan import Plan
    g = load("schemes_1d.npz")
    _, sg, bc = key.split("_")
    grid = g[key + "_grid"]
    N = grid.shape[0]
    _say(f"[gpu] flux1d {key} {solver}")
    plan = Plan(1, sg, solver, "euler", [dict(n0=N, boundary=bc, dx=1 / 32, gamma=1.4, cfl=.5, t_end=1.)],
                varmask=0x13)
    plan.load([grid])
    fl = plan.evolve_space(0)
    plan.close()
    s = O.Scheme(dimension=1, subgrid=sg, solver=solver, boundary=bc, gamma=1.4, dx=1 / 32)
    d = O.reconstruct(grid, s, 0)
    lm, _ = O.local_max_eig(d["jac_state"], s, 0)
    F_or = O.riemann(s, 0, lm, None, d)
    F_ref = g[key + f"_flux_{solver}"]
    got = fl[0]["flux"]
    scale = np.max(np.abs(F_ref), axis=0) + 1e-300
    assert np.max(np.abs(got - F_or) / scale) <= 1e-14
    assert np.max(np.abs(got - F_ref) / scale) <= 1e-13
    assert abs(fl[0]["eigmax"] - float(g[key + "_eigmax"])) <= 1e-14 * float(g[key + "_eigmax"])


@pytest.mark.parametrize("key", SCHEME_KEYS)
def test_lax_wendroff_interface_flux_1d(key):
    """Lax-Wendroff with the sound-speed coefficient: ah_evolve_space fluxes
    equal the oracle's (closed-form radius, same expression order) to the
    last bit on the reference's 1D scheme inputs."""
    from astrea_amd.plan import Plan
    g = load("schemes_1d.npz")
    _, sg, bc = key.split("_")
    grid = g[key + "_grid"]
    N = grid.shape[0]
    _say(f"[gpu] lax-wendroff flux1d {key}")
    plan = Plan(1, sg, "lw", "euler", [dict(n0=N, boundary=bc, dx=1 / 32, gamma=1.4, cfl=.5, t_end=1.)],
                varmask=0x13)
    plan.load([grid])
    fl = plan.evolve_space(0)
    plan.close()
    s = O.Scheme(dimension=1, subgrid=sg, solver="lw", boundary=bc, gamma=1.4, dx=1 / 32)
    d = O.reconstruct(grid, s, 0)
    lm, _ = O.local_max_eig(d["jac_state"], s, 0)
    assert np.array_equal(fl[0]["flux"], O.riemann(s, 0, lm, None, d))
    assert fl[0]["eigmax"] == float(np.max(np.maximum(lm[:-1], lm[1:])))


@pytest.mark.parametrize("dims,subgrid,bc", [(1, "plm", "edge"), (1, "ppm", "wrap"), (1, "weno5", "edge"),
                                             (1, "pcm", "wrap"), (2, "ppm", "wrap"), (2, "plm", "edge")])
def test_lax_wendroff_trajectory_vs_oracle(dims, subgrid, bc):
    """Lax-Wendroff runs (1D Riemann problem 64 cells, 2D quadrants 16^2,
    SSPRK(3,3)) against the oracle's sound-speed form: same steps, bitwise."""
    from astrea_amd.plan import Plan
    from astrea_amd import ic
    import torch
    if dims == 1:
        n, steps = 64, 12
        w = ic.riemann_1d(n, (1., .3, 1.), (.25, .1, .2))
        prob = dict(n0=n, boundary=bc, dx=1 / n, gamma=1.4, cfl=.4, t_end=1e9, max_steps=steps)
        vm = 0x13
    else:
        n, steps = 16, 4
        w = ic.quadrants(n, [(1, .2, .1, 1), (.5, -.3, .2, .4), (.2, .1, -.5, .2), (.8, 0, 0, .9)])
        prob = dict(n0=n, n1=n, boundary=bc, dx=1 / n, gamma=1.4, cfl=.3, t_end=1e9, max_steps=steps)
        vm = 0x17
    q0 = ic.initial_average(torch.from_numpy(w), 1.4, bc, subgrid in ("ppm", "weno5")).numpy()
    _say(f"[gpu] lax-wendroff trajectory {dims}D {subgrid} {bc} n={n} steps={steps}")
    plan = Plan(dims, subgrid, "lw", "ssprk(3,3)", [prob], varmask=vm)
    plan.load([q0])
    plan.reset()
    st = plan.run(chunk=steps)[0]
    out = plan.unpack(0)
    plan.close()
    sch = O.Scheme(dimension=dims, subgrid=subgrid, solver="lw", timestep="ssprk(3,3)", boundary=bc, gamma=1.4,
                   cfl=prob["cfl"], dx=1 / n)
    ref = O.run(q0, sch, 1e9, max_steps=steps)
    assert st["steps"] == steps == ref.steps and not st["failed"]
    e = np.max(rel_l1(out, ref.grid))
    _say(f"[gpu]   rel L1 vs oracle {e:.2e}")
    assert np.array_equal(out, ref.grid)


DOTS_KEYS = [f"{p}_{sg}_{bc}" for p in ("hd", "mhd") for sg in ("pcm", "plm", "ppm", "weno3", "weno5", "weno7")
             for bc in ("edge", "wrap")]


@pytest.mark.parametrize("key", DOTS_KEYS)
def test_dots_interface_flux_1d(key):
    """Osher-Solomon interface fluxes (solvers.py:236-284) against the
    reference's (schemes_1d.npz *_flux_os) and the oracle.  The device takes
    the pseudo-inverse from a Jacobi SVD where the reference calls LAPACK, so
    agreement is to rounding times the conditioning of the eigenvector
    matrices, not bit for bit."""
    from astrea_amd.plan import Plan
    g = load("schemes_1d.npz")
    phys, sg, bc = key.split("_")
    mhd = phys == "mhd"
    gamma = 5 / 3 if mhd else 1.4
    grid = g[key + "_grid"]
    N = grid.shape[0]
    _say(f"[gpu] dots flux1d {key}")
    plan = Plan(1, sg, "os", "euler", [dict(n0=N, boundary=bc, dx=1 / 32, gamma=gamma, cfl=.5, t_end=1.)],
                varmask=0xFF if mhd else 0x13)
    plan.load([grid])
    fl = plan.evolve_space(0)
    plan.close()
    s = O.Scheme(dimension=1, subgrid=sg, solver="os", boundary=bc, gamma=gamma, dx=1 / 32)
    d = O.reconstruct(grid, s, 0)
    F_or = O.riemann(s, 0, None, None, d)
    F_ref = g[key + "_flux_os"]
    got = fl[0]["flux"]
    # per-component scale, floored at 1e-4 of the largest flux: the normal-field
    # flux is identically zero up to rounding noise (~1e-17) on both sides
    scale = np.maximum(np.max(np.abs(F_ref), axis=0), 1e-4 * np.max(np.abs(F_ref)))
    e_or, e_ref = np.max(np.abs(got - F_or) / scale), np.max(np.abs(got - F_ref) / scale)
    _say(f"[gpu]   max scaled error vs oracle {e_or:.2e}, vs reference {e_ref:.2e}")
    assert e_or <= 1e-11 and e_ref <= 1e-11


@pytest.mark.parametrize("key", DOTS_KEYS)
def test_es_interface_flux_1d(key):
    """Entropy-stable interface fluxes (solvers.py:288-391) against the
    reference's (schemes_1d.npz *_flux_es) and the oracle: the same entries
    are non-finite (most of them, the reference's averages being
    half-differences), the finite ones agree to rounding (the reference's
    8x8 products go through BLAS)."""
    from astrea_amd.plan import Plan
    g = load("schemes_1d.npz")
    phys, sg, bc = key.split("_")
    gamma = 5 / 3 if phys == "mhd" else 1.4
    grid = g[key + "_grid"]
    N = grid.shape[0]
    _say(f"[gpu] es flux1d {key}")
    plan = Plan(1, sg, "es", "euler", [dict(n0=N, boundary=bc, dx=1 / 32, gamma=gamma, cfl=.5, t_end=1.)],
                varmask=0xFF)
    plan.load([grid])
    fl = plan.evolve_space(0)
    plan.close()
    s = O.Scheme(dimension=1, subgrid=sg, solver="es", boundary=bc, gamma=gamma, dx=1 / 32)
    d = O.reconstruct(grid, s, 0)
    with np.errstate(all="ignore"):
        F_or = O.riemann(s, 0, None, None, d)
    F_ref = g[key + "_flux_es"]
    got = fl[0]["flux"]
    assert np.array_equal(np.isfinite(got), np.isfinite(F_ref))
    assert np.array_equal(np.isfinite(F_or), np.isfinite(F_ref))
    fin = np.isfinite(F_ref)
    big = np.max(np.abs(np.where(fin, F_ref, 0)))
    scale = np.maximum(np.max(np.abs(np.where(fin, F_ref, 0)), axis=0), 1e-4 * big)
    e_or = np.max(np.where(fin, np.abs(got - np.where(fin, F_or, 0)) / scale, 0))
    e_ref = np.max(np.where(fin, np.abs(got - np.where(fin, F_ref, 0)) / scale, 0))
    _say(f"[gpu]   {int((~fin).sum())} of {fin.size} non-finite in both; finite max scaled error vs oracle "
         f"{e_or:.2e}, vs reference {e_ref:.2e}")
    assert e_or <= 1e-11 and e_ref <= 1e-11


def test_es_run_fails_where_the_reference_does():
    """A 1D run with the entropy-stable flux: the state turns non-finite in
    the oracle after the first step; the device raises the failure flag at
    the next stage-0 spectral radius (where the reference's LAPACK call would
    raise) and keeps the last finite state."""
    from astrea_amd.plan import Plan
    from astrea_amd import ic
    import torch
    n = 64
    w = ic.riemann_1d(n, (1., 0., 1.), (.125, 0., .1))
    q0 = ic.initial_average(torch.from_numpy(w), 1.4, "edge", False).numpy()
    _say("[gpu] es 1D sod 64 run")
    plan = Plan(1, "plm", "es", "ssprk(3,3)", [dict(n0=n, boundary="edge", dx=1 / n, gamma=1.4, cfl=.5, t_end=1e9,
                                                     max_steps=5)], varmask=0xFF)
    plan.load([q0])
    plan.reset()
    st = plan.run(chunk=5)[0]
    out = plan.unpack(0)
    plan.close()
    s = O.Scheme(dimension=1, subgrid="plm", solver="es", timestep="ssprk(3,3)", boundary="edge", gamma=1.4, cfl=.5,
                 dx=1 / n)
    first_bad = None
    with np.errstate(all="ignore"):
        for k in range(1, 6):
            r = O.run(q0, s, 1e9, max_steps=k)
            if not np.isfinite(r.grid).all():
                first_bad = k
                break
    _say(f"[gpu]   oracle non-finite after step {first_bad}; device steps {st['steps']} failed {st['failed']}")
    assert first_bad is not None
    assert st["failed"] and st["steps"] == first_bad - 1
    assert np.isfinite(out).all()
    if first_bad > 1:
        with np.errstate(all="ignore"):
            ok = O.run(q0, s, 1e9, max_steps=first_bad - 1)
        assert np.max(rel_l1(out, ok.grid)) <= 1e-12
    else:
        assert np.array_equal(out, q0)


@pytest.mark.parametrize("dims,subgrid", [(1, "plm"), (1, "ppm"), (2, "plm")])
def test_dots_trajectory_vs_oracle(dims, subgrid):
    """Osher-Solomon runs (1D Riemann problem 64 cells, 2D quadrants 16^2),
    SSPRK(3,3), against the oracle: same steps, rel L1 per variable."""
    from astrea_amd.plan import Plan
    from astrea_amd import ic
    import torch
    if dims == 1:
        n, steps = 64, 20
        w = ic.riemann_1d(n, (2.5, .7, 4.), (.3, -.4, .2))
        prob = dict(n0=n, boundary="edge", dx=1 / n, gamma=1.4, cfl=.5, t_end=1e9, max_steps=steps)
        vm = 0x13
    else:
        n, steps = 16, 4
        w = ic.quadrants(n, [(1, .2, .1, 1), (.5, -.3, .2, .4), (.2, .1, -.5, .2), (.8, 0, 0, .9)])
        prob = dict(n0=n, n1=n, boundary="edge", dx=1 / n, gamma=1.4, cfl=.4, t_end=1e9, max_steps=steps)
        vm = 0x17
    q0 = ic.initial_average(torch.from_numpy(w), 1.4, "edge", subgrid == "ppm").numpy()
    _say(f"[gpu] dots trajectory {dims}D {subgrid} n={n} steps={steps}")
    plan = Plan(dims, subgrid, "os", "ssprk(3,3)", [prob], varmask=vm)
    plan.load([q0])
    plan.reset()
    st = plan.run(chunk=steps)[0]
    out = plan.unpack(0)
    plan.close()
    sch = O.Scheme(dimension=dims, subgrid=subgrid, solver="os", timestep="ssprk(3,3)", boundary="edge", gamma=1.4,
                   cfl=prob["cfl"], dx=1 / n)
    ref = O.run(q0, sch, 1e9, max_steps=steps)
    assert st["steps"] == steps == ref.steps and not st["failed"]
    e = np.max(rel_l1(out, ref.grid))
    _say(f"[gpu]   rel L1 vs oracle {e:.2e}")
    assert e <= 1e-12


@pytest.mark.parametrize("case", [("ppm", "hllc", "wrap"), ("plm", "hllc", "edge"), ("ppm", "llf", "edge")])
@pytest.mark.parametrize("parity", [0, 1])
def test_interface_flux_2d(case, parity):
    """2D double solve + transverse correction (space_2d.npz) at both step
    parities (the sweep-order flip changes the HLLC normal index)."""
    from astrea_amd.plan import Plan
    subgrid, solver, bc = case
    g = load("space_2d.npz")
    key = f"{subgrid}_{solver}_{bc}_p{parity}"
    grid = g[key + "_grid"]
    N = grid.shape[0]
    _say(f"[gpu] flux2d {key}")
    plan = Plan(2, subgrid, solver, "euler", [dict(n0=N, n1=N, boundary=bc, dx=1 / 12, gamma=1.4, cfl=.4, t_end=1.)],
                varmask=0x17)
    plan.load([grid])
    fl = plan.evolve_space(parity)
    plan.close()
    for ax in (0, 1):
        ref = g[key + f"_flux_ax{ax}"]
        scale = np.max(np.abs(ref), axis=(0, 1)) + 1e-300
        err = np.max(np.abs(fl[ax]["flux"] - ref) / scale)
        _say(f"[gpu]   axis {ax}: max scaled flux error {err:.2e}")
        assert err <= 1e-12
        e_ref = float(g[key + f"_eigmax_ax{ax}"])
        assert abs(fl[ax]["eigmax"] - e_ref) <= 1e-14 * e_ref


def test_failure_flag_keeps_last_good_state():
    """The seeded Riemann problem on which the reference raises in step
    steps_ok + 1: the device reports failure after steps_ok good steps and
    keeps that state."""
    m = meta()["cases"]["rp1d_fail"]
    g = load("steps.npz")
    grid0 = g["rp1d_fail_grid0"]
    N = grid0.shape[0]
    mm = dict(dimension=1, subgrid="ppm", solver="hllc", timestep="ssprk(3,3)", boundary="edge", gamma=1.4, cfl=.5,
              dx=1 / N, t_end=m["t_end"], checkpoints=1, max_steps=None, magnetic_2d=False)
    _say("[gpu] failure flag rp1d_fail")
    out, st = gpu_run(grid0, mm)
    assert st["failed"] and not st["active"]
    assert st["steps"] == m["steps_ok"]
    ref = oracle_run(grid0, dict(mm, max_steps=m["steps_ok"]))
    assert np.max(rel_l1(out, ref.grid)) <= 1e-12


@pytest.mark.parametrize("subgrid,solver", [("ppm", "hllc"), ("plm", "llf"), ("weno5", "hllc")])
def test_conservation_periodic_2d(subgrid, solver):
    rng = np.random.default_rng(7)
    N = 32
    c = (np.arange(N) + .5) / N
    x, y = np.meshgrid(c, c, indexing="ij")
    w = np.zeros((N, N, 8))
    w[..., 0] = 1 + .3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    w[..., 1] = .2 * np.cos(2 * np.pi * y) + .01 * rng.uniform(-1, 1, (N, N))
    w[..., 2] = .1 * np.sin(2 * np.pi * x)
    w[..., 4] = 1.5
    s = O.Scheme(dimension=2, subgrid=subgrid, solver=solver, boundary="wrap", dx=1 / N, cfl=.4)
    q0 = O.initial_average(w, s)
    m = dict(dimension=2, subgrid=subgrid, solver=solver, timestep="ssprk(3,3)", boundary="wrap", gamma=1.4,
             cfl=.4, dx=1 / N, t_end=10., checkpoints=1, max_steps=10, magnetic_2d=False)
    _say(f"[gpu] conservation {subgrid} {solver}")
    out, st = gpu_run(q0, m)
    assert st["steps"] == 10
    tot0, tot1 = q0.sum((0, 1)), out.sum((0, 1))
    assert np.all(np.abs(tot1 - tot0) <= 1e-13 * np.maximum(1, np.abs(tot0)))


@pytest.mark.parametrize("n,subgrid,bc", [(17, "ppm", "wrap"), (17, "plm", "edge"), (33, "ppm", "edge"),
                                          (25, "weno5", "wrap"), (13, "ppm", "edge"), (65, "plm", "wrap")])
def test_ragged_tiles_2d(n, subgrid, bc):
    """Grids whose last reconstruction and flux tiles hold a single column
    or row (16k+1, 8k+1, 4k+1 cells; tiles of 32x8/64x4 cells and 16x12
    interfaces) or a few: device FP64 equals the oracle after 6 SSPRK(3,3)
    steps."""
    rng = np.random.default_rng(n)
    c = (np.arange(n) + .5) / n
    x, y = np.meshgrid(c, c, indexing="ij")
    w = np.zeros((n, n, 8))
    w[..., 0] = 1 + .3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + .02 * rng.uniform(-1, 1, (n, n))
    w[..., 1] = .2 * np.cos(2 * np.pi * y)
    w[..., 2] = .1 * np.sin(2 * np.pi * x) + .01 * rng.uniform(-1, 1, (n, n))
    w[..., 4] = 1.5 + .1 * (x > .5)
    s = O.Scheme(dimension=2, subgrid=subgrid, solver="hllc", boundary=bc, dx=1 / n, cfl=.4)
    q0 = O.initial_average(w, s)
    m = dict(dimension=2, subgrid=subgrid, solver="hllc", timestep="ssprk(3,3)", boundary=bc, gamma=1.4,
             cfl=.4, dx=1 / n, t_end=10., checkpoints=1, max_steps=6, magnetic_2d=False)
    _say(f"[gpu] ragged tiles {n}^2 {subgrid} {bc}")
    out, st = gpu_run(q0, m)
    ref = oracle_run(q0, m)
    assert st["steps"] == ref.steps == 6 and not st["failed"]
    e = np.max(rel_l1(out, ref.grid))
    _say(f"[gpu]   rel L1 vs oracle {e:.2e}")
    assert e <= 1e-12


def test_mhd_ragged_tiles():
    """Orszag-Tang on 17^2 (single-column last tiles in every kernel of the
    constrained-transport path), WENO5 + HLLD, 4 steps: device vs oracle."""
    from astrea_amd.plan import Plan
    from astrea_amd import ic
    import torch
    n = 17
    q0 = ic.initial_average(torch.from_numpy(ic.orszag_tang(n)), 5 / 3, "wrap", True).numpy()
    _say(f"[gpu] mhd ragged tiles {n}^2")
    plan = Plan(2, "weno5", "hlld", "ssprk(3,3)", [dict(n0=n, n1=n, boundary="wrap", dx=1 / n, gamma=5 / 3, cfl=.4,
                                                      t_end=1e9, max_steps=4)], varmask=0x77, magnetic_2d=True)
    plan.load([q0])
    plan.reset()
    st = plan.run(chunk=4)[0]
    out = plan.unpack(0)
    plan.close()
    s = O.Scheme(dimension=2, subgrid="weno5", solver="hlld", boundary="wrap", gamma=5 / 3, cfl=.4, dx=1 / n,
                 magnetic_2d=True)
    ref = O.run(q0, s, 1e9, max_steps=4)
    assert st["steps"] == ref.steps == 4 and not st["failed"]
    e = np.max(rel_l1(out, ref.grid))
    _say(f"[gpu]   rel L1 vs oracle {e:.2e}")
    assert e <= 1e-12


@pytest.mark.parametrize("config,subgrid,solver,steps", [("orszag-tang", "weno5", "hllc", 4),
                                                          ("orszag-tang", "plm", "hllc", 4),
                                                          ("mhd rotor", "ppm", "hllc", 3),
                                                          ("orszag-tang", "plm", "lw", 4),
                                                          ("mhd rotor", "weno5", "lw", 3)])
def test_mhd_2d_hllc_and_lw(config, subgrid, solver, steps):
    """2D MHD with constrained transport under HLLC (a magnetic-2D
    configuration keeps 'hllc': generic.py:266-269 resets it for 1D magnetic
    configurations only) and Lax-Wendroff (sound-speed form), 16^2 from the
    configuration's own initial state: device vs oracle.  The corner speeds
    follow the solver category (mag_field.py:135-159): the HLL estimate for
    HLLC, the Jacobian's extreme eigenvalues for Lax-Wendroff."""
    from astrea_amd import ic
    from astrea_amd.options import handle_variables
    from astrea_amd.plan import Plan
    sv = handle_variables({"config": config, "cells": 16, "solver": solver, "subgrid": subgrid,
                           "timestep": "ssprk(3,3)", "dimension": 2})
    q0 = np.asarray(ic.initialise(sv))
    n = q0.shape[0]
    _say(f"[gpu] 2D MHD {config} {subgrid}/{solver} {n}^2, {steps} steps")
    plan = Plan(2, subgrid, solver, "ssprk(3,3)", [dict(n0=n, n1=n, boundary=sv.boundary, dx=sv.dx, gamma=sv.gamma,
                                                        cfl=sv.cfl, t_end=1e9, max_steps=steps)], varmask=0x77,
                magnetic_2d=True, solver_category=sv.solver_category)
    plan.load([q0])
    plan.reset()
    st = plan.run(chunk=steps)[0]
    out = plan.unpack(0)
    plan.close()
    s = O.Scheme(dimension=2, subgrid=subgrid, solver=solver, boundary=sv.boundary, gamma=sv.gamma, cfl=sv.cfl,
                 dx=sv.dx, magnetic_2d=True)
    with np.errstate(all="ignore"):
        ref = O.run(q0, s, 1e9, max_steps=steps)
    assert st["steps"] == ref.steps == steps and not st["failed"]
    e = np.max(rel_l1(out, ref.grid))
    _say(f"[gpu]   rel L1 vs oracle {e:.2e}")
    assert e <= 1e-12


@pytest.mark.parametrize("subgrid,solver,bc", [("weno5", "hlld", "wrap"), ("plm", "llf", "edge")])
def test_mhd_ssprk104_trajectory(subgrid, solver, bc):
    """SSPRK(10,4) with constrained transport: its accumulating stages add a
    refined increment, k += refine_grid(dt/6 L) (evolvers.py:89-90, 97-99),
    the others refine the combined register; Orszag-Tang 16^2, 3 steps,
    device vs oracle."""
    from astrea_amd.plan import Plan
    from astrea_amd import ic
    import torch
    n = 16
    q0 = ic.initial_average(torch.from_numpy(ic.orszag_tang(n)), 5 / 3, bc, subgrid == "weno5").numpy()
    _say(f"[gpu] mhd ssprk(10,4) {n}^2 {subgrid}/{solver}/{bc}")
    plan = Plan(2, subgrid, solver, "ssprk(10,4)", [dict(n0=n, n1=n, boundary=bc, dx=1 / n, gamma=5 / 3, cfl=.4,
                                                         t_end=1e9, max_steps=3)], varmask=0x77, magnetic_2d=True)
    plan.load([q0])
    plan.reset()
    st = plan.run(chunk=3)[0]
    out = plan.unpack(0)
    plan.close()
    s = O.Scheme(dimension=2, subgrid=subgrid, solver=solver, timestep="ssprk(10,4)", boundary=bc, gamma=5 / 3,
                 cfl=.4, dx=1 / n, magnetic_2d=True)
    with np.errstate(all="ignore"):
        ref = O.run(q0, s, 1e9, max_steps=3)
    assert st["steps"] == ref.steps == 3 and not st["failed"]
    e = np.max(rel_l1(out, ref.grid))
    _say(f"[gpu]   rel L1 vs oracle {e:.2e}")
    assert e <= 1e-12


@pytest.mark.parametrize("n,subgrid,solver", [(17, "weno5", "hlld"), (24, "ppm", "hlld"), (20, "plm", "llf")])
def test_mhd_edge_boundary(n, subgrid, solver):
    """The constrained-transport path with edge (outlet) boundaries, whose
    padded intermediates are clamped rather than wrapped: Orszag-Tang data,
    5 SSPRK(3,3) steps, device vs oracle."""
    from astrea_amd.plan import Plan
    from astrea_amd import ic
    import torch
    q0 = ic.initial_average(torch.from_numpy(ic.orszag_tang(n)), 5 / 3, "edge", True).numpy()
    _say(f"[gpu] mhd edge boundary {n}^2 {subgrid} {solver}")
    plan = Plan(2, subgrid, solver, "ssprk(3,3)", [dict(n0=n, n1=n, boundary="edge", dx=1 / n, gamma=5 / 3,
                                                        cfl=.4, t_end=1e9, max_steps=5)], varmask=0x77,
                magnetic_2d=True)
    plan.load([q0])
    plan.reset()
    st = plan.run(chunk=5)[0]
    out = plan.unpack(0)
    plan.close()
    s = O.Scheme(dimension=2, subgrid=subgrid, solver=solver, boundary="edge", gamma=5 / 3, cfl=.4, dx=1 / n,
                 magnetic_2d=True)
    with np.errstate(all="ignore"):
        ref = O.run(q0, s, 1e9, max_steps=5)
    assert st["steps"] == ref.steps == 5 and not st["failed"]
    e = np.max(rel_l1(out, ref.grid))
    _say(f"[gpu]   rel L1 vs oracle {e:.2e}")
    assert e <= 1e-12


@pytest.mark.parametrize("bc", ["edge", "wrap"])
@pytest.mark.parametrize("high_order", [False, True])
def test_device_initial_state_equals_host_pipeline(bc, high_order):
    """ah_plan_init_ic builds the conservative initial averages on the device
    bit for bit as the host pipeline does (itself pinned to the oracle's
    constructor restatement): 1D Riemann problems and 2D four-quadrant
    problems of several sizes, non-square grids included, packed in one plan."""
    from astrea_amd.plan import Plan
    from astrea_amd import ic
    import torch
    sub = "ppm" if high_order else "plm"
    _say(f"[gpu] device initial state, {bc}, high_order={high_order}")
    r1 = [(37, (2.5, .7, 4.), (.3, -.4, .2)), (64, (1., 0., 1.), (.125, 0., .1)), (5, (1., .5, 2.), (3., -1., .5))]
    plan = Plan(1, sub, "hllc", "ssprk(3,3)", [dict(n0=n, boundary=bc, dx=1 / n, gamma=1.4, cfl=.5, t_end=.1)
                                               for n, _, _ in r1], varmask=0x13)
    plan.load([ic.initial_average(torch.from_numpy(ic.riemann_1d(n, l, r)), 1.4, bc, high_order).numpy()
               for n, l, r in r1])
    host = plan.state.clone()
    plan.state.zero_()
    for i, (n, l, r) in enumerate(r1):
        plan.init_riemann(i, l, r)
    torch.cuda.synchronize()
    assert torch.equal(plan.state, host)
    plan.close()
    st = [(1, .2, .1, 1), (.5, -.3, .2, .4), (.2, .1, -.5, .2), (.8, 0, 0, .9)]
    r2 = [(21, 21), (64, 64), (24, 17), (9, 33)]
    plan = Plan(2, sub, "hllc", "ssprk(3,3)", [dict(n0=a, n1=b, boundary=bc, dx=1 / a, gamma=1.4, cfl=.4, t_end=.1)
                                               for a, b in r2], varmask=0x17)
    grids = []
    for a, b in r2:
        c0, c1 = ic.cell_centres(a), ic.cell_centres(b)
        x, y = np.meshgrid(c0, c1, indexing="ij")
        w = np.zeros((a, b, 8))
        masks = [(x <= .5) & (y < .5), (x > .5) & (y < .5), (x <= .5) & (y >= .5), (x > .5) & (y >= .5)]
        for m, (rr, u, v, p) in zip(masks, st):
            w[m, 0], w[m, 1], w[m, 2], w[m, 4] = rr, u, v, p
        grids.append(ic.initial_average(torch.from_numpy(w), 1.4, bc, high_order).numpy())
    plan.load(grids)
    host = plan.state.clone()
    plan.state.zero_()
    for i in range(len(r2)):
        plan.init_quadrants(i, st)
    torch.cuda.synchronize()
    assert torch.equal(plan.state, host)
    plan.close()


def _shear(n, seed=20261202):
    from astrea_amd import ic
    import torch
    w = ic.shear_layer(n, seed)
    return ic.initial_average(torch.from_numpy(w), 1.4, "wrap", True).numpy()


@pytest.mark.parametrize("subgrid,flattener", [("ppm", "astrea"), ("ppm", "colella90"), ("plm", "off")])
def test_shear_layer_physical_compat(subgrid, flattener):
    """The benchmark problem at 32x32 with the physical solver axis and the
    PPM flattener variants: device FP64 equals the oracle."""
    from astrea_amd.plan import Plan
    n, steps = 32, 12
    q0 = _shear(n)
    _say(f"[gpu] shear layer {n}^2 {subgrid} flattener={flattener} compat=physical")
    plan = Plan(2, subgrid, "hllc", "ssprk(3,3)", [dict(n0=n, n1=n, boundary="wrap", dx=1 / n, gamma=1.4, cfl=.4,
                                                        t_end=1e9, max_steps=steps)],
                flattener=flattener, varmask=0x17, compat="physical")
    plan.load([q0])
    plan.reset()
    st = plan.run(chunk=steps)[0]
    out = plan.unpack(0)
    plan.close()
    s = O.Scheme(dimension=2, subgrid=subgrid, solver="hllc", boundary="wrap", gamma=1.4, cfl=.4, dx=1 / n,
                 flattener=flattener, compat="physical")
    ref = O.run(q0, s, 1e9, max_steps=steps)
    assert st["steps"] == steps and not st["failed"]
    e = np.max(rel_l1(out, ref.grid))
    _say(f"[gpu]   rel L1 vs oracle {e:.2e}")
    assert e <= 1e-12


def _split_vs_fp64(dims, recon, prob, q0, steps=None, flattener="off", varmask=0x17, solver="hllc",
                   magnetic_2d=False):
    """Run the FP64 and split paths from one initial state; rel L1 of split
    against FP64 (max over variables) and both statuses."""
    from astrea_amd.plan import Plan
    outs, sts = {}, {}
    pr = dict(prob, max_steps=steps) if steps else prob
    for prec in ("fp64", "split"):
        plan = Plan(dims, recon, solver, "ssprk(3,3)", [pr], flattener=flattener, varmask=varmask,
                    compat="physical", precision=prec, magnetic_2d=magnetic_2d)
        plan.load([q0])
        plan.reset()
        sts[prec] = plan.run(chunk=64)[0]
        outs[prec] = plan.unpack(0)
        plan.close()
    return float(np.max(rel_l1(outs["split"], outs["fp64"]))), sts


def test_split_precision_shear_layer():
    """Split path (FP32 Riemann-solver coefficients and spectral radius, FP64
    values) on the shear layer with the flattener, 10 steps: rel L1 against
    FP64 <= 1e-5.  (Over ~50 steps this problem amplifies a 1e-15 relative
    perturbation of the initial state to ~1e-3 in FP64 itself, through the
    flattener's discontinuous switches, so longer runs bound nothing.)"""
    n = 128
    _say(f"[gpu] split vs fp64: shear layer {n}^2, flattener, 10 steps")
    e, st = _split_vs_fp64(2, "ppm", dict(n0=n, n1=n, boundary="wrap", dx=1 / n, gamma=1.4, cfl=.4, t_end=1e9),
                           _shear(n), steps=10, flattener="astrea")
    _say(f"[gpu]   rel L1 {e:.2e}")
    assert st["split"]["steps"] == 10 and not st["split"]["failed"]
    assert e <= 1e-5


def test_split_precision_orszag_tang():
    """Split path of 2D MHD (FP32 spectral radius and primitive-recovery
    reciprocal; HLLD, constrained transport and values FP64) on the
    Orszag-Tang vortex, 128^2, 12 steps: rel L1 against FP64 <= 1e-5."""
    from astrea_amd import ic
    n = 128
    _say(f"[gpu] split vs fp64: Orszag-Tang {n}^2, WENO5/HLLD/CT, 12 steps")
    import torch
    q0 = ic.initial_average(torch.from_numpy(ic.orszag_tang(n)), 5 / 3, "wrap", True).numpy()
    e, st = _split_vs_fp64(2, "weno5", dict(n0=n, n1=n, boundary="wrap", dx=1 / n, gamma=5 / 3, cfl=.4,
                                            t_end=1e9), q0, steps=12, varmask=0x77, solver="hlld",
                           magnetic_2d=True)
    _say(f"[gpu]   rel L1 {e:.2e}")
    assert st["split"]["steps"] == 12 and not st["split"]["failed"]
    assert e <= 1e-5


def test_split_precision_brio_wu():
    """Split path of 1D MHD with HLLD on Brio-Wu, 400 cells, 40 steps: rel
    L1 against FP64 <= 1e-5 (with FP32 fast speeds inside HLLD this case
    left FP64 by 0.1 at the discontinuity after six steps: a region switch
    of the reference's HLLD flipped, so HLLD stays FP64 on the split path)."""
    from astrea_amd import ic
    from astrea_amd.options import handle_variables
    sv = handle_variables({"config": "brio-wu", "cells": 400, "solver": "hlld", "subgrid": "plm",
                           "timestep": "ssprk(3,3)"}, t_end=.1)
    g = np.asarray(ic.initialise(sv))
    _say("[gpu] split vs fp64: Brio-Wu 400 cells, PLM/HLLD, 40 steps")
    e, st = _split_vs_fp64(1, "plm", dict(n0=400, boundary=sv.boundary, dx=sv.dx, gamma=sv.gamma, cfl=.5,
                                          t_end=1.), g, steps=40, varmask=0xFF, solver="hlld")
    _say(f"[gpu]   steps {st['split']['steps']} failed {st['fp64']['failed']}/{st['split']['failed']} rel L1 {e:.2e}")
    assert st["fp64"]["steps"] == st["split"]["steps"] == 40 and not st["split"]["failed"]
    assert e <= 1e-5


def test_split_precision_quadrants():
    """Four-quadrant requests of the served stream (PLM and PPM, 128^2), each
    to t_end = 0.25: split path rel L1 against FP64 <= 1e-5 on every request
    the FP64 path completes; a request whose FP64 solve fails (the scheme
    reaches negative density and then a non-finite state) fails on the split
    path too."""
    from astrea_amd import ic
    import torch
    compared = 0
    for k, sp in enumerate(ic.quadrant_requests(6, 11, (128,))):
        w = ic.quadrants(sp["n"], sp["states"])
        q0 = ic.initial_average(torch.from_numpy(w), 1.4, "edge", sp["subgrid"] == "ppm").numpy()
        n = sp["n"]
        _say(f"[gpu] split vs fp64: quadrant request {k} {sp['subgrid']} {n}^2 to t=0.25")
        e, st = _split_vs_fp64(2, sp["subgrid"], dict(n0=n, n1=n, boundary="edge", dx=1 / n, gamma=1.4, cfl=.4,
                                                      t_end=.25), q0)
        _say(f"[gpu]   steps {st['fp64']['steps']}/{st['split']['steps']} failed {st['fp64']['failed']}/"
             f"{st['split']['failed']} rel L1 {e:.2e}")
        assert st["fp64"]["failed"] == st["split"]["failed"]
        if st["fp64"]["failed"]:
            continue
        assert st["fp64"]["steps"] == st["split"]["steps"]
        assert e <= 1e-5
        compared += 1
    assert compared >= 3


def test_split_precision_riemann_ensemble():
    """1D Riemann problems of the ensemble (PPM/HLLC, N in [256, 1024]), each
    to its own t_end: split path rel L1 against FP64 <= 1e-5 on every problem
    the FP64 path completes; FP64 failures fail on the split path too."""
    grids, probs = _ensemble(6, 5, 256, 1024, t_scale=1.0)
    compared = 0
    for k, (g, p) in enumerate(zip(grids, probs)):
        _say(f"[gpu] split vs fp64: ensemble problem {k} N={p['n0']} to t={p['t_end']:.3f}")
        e, st = _split_vs_fp64(1, "ppm", p, g, varmask=0x13)
        _say(f"[gpu]   steps {st['fp64']['steps']}/{st['split']['steps']} failed {st['fp64']['failed']}/"
             f"{st['split']['failed']} rel L1 {e:.2e}")
        assert st["fp64"]["failed"] == st["split"]["failed"]
        if st["fp64"]["failed"]:
            continue
        assert st["fp64"]["steps"] == st["split"]["steps"]
        assert e <= 1e-5
        compared += 1
    assert compared >= 3


def _ensemble(nprob, seed, nmin, nmax, t_scale=1 / 8):
    from astrea_amd import ic
    import torch
    specs = ic.ensemble_1d(nprob, seed, nmin, nmax)
    grids, probs = [], []
    for sp in specs:
        w = ic.riemann_1d(sp["n"], sp["left"], sp["right"])
        grids.append(ic.initial_average(torch.from_numpy(w), 1.4, "edge", True).numpy())
        probs.append(dict(n0=sp["n"], boundary="edge", dx=1 / sp["n"], gamma=1.4, cfl=.5, t_end=sp["t_end"] * t_scale))
    return grids, probs


@pytest.mark.parametrize("use_graph,compact_every", [(False, 1), (True, 1), (True, 3)])
def test_batch_equals_alone(use_graph, compact_every):
    """A ragged batch of seeded 1D Riemann problems (PPM, HLLC, SSPRK(3,3),
    CFL 0.5) in one plan gives every problem the same result, step count and
    failure status as running it alone (bitwise), with and without the
    captured-graph replay, compacting the tile lists after every chunk or
    every third."""
    from astrea_amd.plan import Plan
    from astrea_amd import batch
    grids, probs = _ensemble(12, 11, 32, 300)
    _say(f"[gpu] batch of {len(grids)} problems, graph={use_graph}, compact_every={compact_every}, "
         f"sizes={[p['n0'] for p in probs]}")
    plan = Plan(1, "ppm", "hllc", "ssprk(3,3)", probs, varmask=0x13)
    plan.load(grids)
    plan.reset()
    st, secs, nch = batch.run(plan, chunk=8, use_graph=use_graph, compact_every=compact_every)
    outs = plan.unpack()
    plan.close()
    for i, (g, p) in enumerate(zip(grids, probs)):
        alone = Plan(1, "ppm", "hllc", "ssprk(3,3)", [p], varmask=0x13)
        alone.load([g])
        alone.reset()
        s1 = alone.run(chunk=64)[0]
        o1 = alone.unpack(0)
        alone.close()
        assert s1["steps"] == st[i]["steps"] and s1["failed"] == st[i]["failed"], (i, s1, st[i])
        assert np.array_equal(o1, outs[i]), i
    _say(f"[gpu]   {nch} chunks, failed={sum(s['failed'] for s in st)}")


def test_end_of_step_launch_and_fold_agree():
    """A step ends in its last combination launch on plans with at most 64
    combination tiles (256 cells each in 1D) per problem, and in a launch of
    its own on the others.  A batch of the eleven small seeded Riemann
    problems plus a 20 000-cell Sod problem (79 tiles: the plan takes the
    separate launch) gives every small problem the state, step count and
    failure status of running it alone (where the combination ends the step),
    bit for bit; the large problem ends in the oracle's state after its 12
    steps."""
    from astrea_amd.plan import Plan
    from astrea_amd import ic
    import torch
    grids, probs = _ensemble(12, 11, 32, 300)
    n = 20000
    w = ic.riemann_1d(n, (1., 0., 1.), (.125, 0., .1))
    big = ic.initial_average(torch.from_numpy(w), 1.4, "edge", True).numpy()
    grids.append(big)
    probs.append(dict(n0=n, boundary="edge", dx=1 / n, gamma=1.4, cfl=.5, t_end=1e9, max_steps=12))
    _say(f"[gpu] end-of-step launch vs fold: sizes={[p['n0'] for p in probs]}")
    plan = Plan(1, "ppm", "hllc", "ssprk(3,3)", probs, varmask=0x13)
    plan.load(grids)
    plan.reset()
    st = plan.run(chunk=8)
    outs = plan.unpack()
    plan.close()
    for i in range(len(probs) - 1):
        alone = Plan(1, "ppm", "hllc", "ssprk(3,3)", [probs[i]], varmask=0x13)
        alone.load([grids[i]])
        alone.reset()
        s1 = alone.run(chunk=64)[0]
        o1 = alone.unpack(0)
        alone.close()
        assert s1["steps"] == st[i]["steps"] and s1["failed"] == st[i]["failed"], (i, s1, st[i])
        assert np.array_equal(o1, outs[i]), i
    assert st[-1]["steps"] == 12 and not st[-1]["failed"]
    s = O.Scheme(dimension=1, subgrid="ppm", solver="hllc", boundary="edge", gamma=1.4, cfl=.5, dx=1 / n)
    ref = O.run(big, s, 1e9, 1, 12)
    assert ref.steps == 12
    assert np.max(rel_l1(outs[-1], ref.grid)) <= 1e-12


def _totals_conserved(q0, q1):
    """Per-variable totals equal to 1e-13 of the total magnitude; each total
    is a pairwise sum over a contiguous copy (numpy sums strided reductions
    one by one, whose own rounding would exceed the tolerance at 2048^2)."""
    tot = lambda q, c: np.ascontiguousarray(q[..., c]).sum()
    return np.array([abs(tot(q1, c) - tot(q0, c)) <= 1e-13 * max(np.abs(q0[..., c]).sum(), 1e-300)
                     for c in range(q0.shape[-1])])


@pytest.mark.parametrize("case", ["shear1024", "ot2048"])
def test_full_size_periodic_properties(case):
    """At the benchmark sizes (shear layer 1024^2 PPM+flattener/HLLC, and
    Orszag-Tang 2048^2 WENO5/HLLD/constrained transport, both periodic,
    SSPRK(3,3)), two size-independent properties: periodic translation
    equivariance -- the run started from the initial state rolled by
    (37, 101) cells ends in the rolled final state, bit for bit -- and the
    totals of every conserved variable kept to round-off."""
    from astrea_amd import ic
    from astrea_amd.plan import Plan
    import torch
    if case == "shear1024":
        n, steps, gamma = 1024, 12, 1.4
        w = ic.shear_layer(n, 20261202)
        mk = lambda: Plan(2, "ppm", "hllc", "ssprk(3,3)", [prob], flattener="astrea", varmask=0x17,
                          compat="physical")
        high = True
    else:
        n, steps, gamma = 2048, 3, 5 / 3
        w = ic.orszag_tang(n)
        mk = lambda: Plan(2, "weno5", "hlld", "ssprk(3,3)", [prob], varmask=0x77, magnetic_2d=True)
        high = True
    prob = dict(n0=n, n1=n, boundary="wrap", dx=1 / n, gamma=gamma, cfl=.4, t_end=1e9, max_steps=steps)
    q0 = ic.initial_average(torch.from_numpy(w), gamma, "wrap", high).numpy()
    outs = []
    for shift in ((0, 0), (37, 101)):
        _say(f"[gpu] full size {case}: {steps} steps, initial state rolled by {shift}")
        plan = mk()
        plan.load([np.roll(q0, shift, axis=(0, 1))])
        plan.reset()
        st = plan.run(chunk=steps)[0]
        assert st["steps"] == steps and not st["failed"]
        outs.append(plan.unpack(0))
        plan.close()
    assert np.array_equal(np.roll(outs[0], (37, 101), axis=(0, 1)), outs[1])
    # rho, rho v, E; the field totals are not kept by the reference's
    # constrained transport (its cell-centred B is re-interpolated from the
    # face values every stage, mag_field.py:191-210)
    ok = _totals_conserved(q0, outs[0])
    assert np.all(ok[[0, 1, 2, 4]]), ok


def test_full_size_ensemble_batch_equals_alone():
    """The full seeded ensemble of 512 ragged 1D Riemann problems (N in
    [256, 8192], PPM/HLLC/SSPRK(3,3)) batched in one plan for 40 steps; four
    of its problems (the smallest, the largest and two others) run alone give
    the same states and step counts bit for bit."""
    from astrea_amd import workloads
    from astrea_amd.plan import Plan
    inits, probs = workloads.ensemble_inputs(512)
    probs = [dict(p, max_steps=40) for p in probs]
    _say(f"[gpu] full ensemble: {len(probs)} problems, {sum(p['n0'] for p in probs)} cells, 40 steps")
    plan = Plan(1, "ppm", "hllc", "ssprk(3,3)", probs, varmask=0x13)
    workloads._start(plan, inits)
    plan.reset()
    st = plan.run(chunk=40)
    sizes = [p["n0"] for p in probs]
    picks = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 100, 333})
    outs = {i: plan.unpack(i) for i in picks}
    plan.close()
    for i in picks:
        alone = Plan(1, "ppm", "hllc", "ssprk(3,3)", [probs[i]], varmask=0x13)
        workloads._start(alone, [inits[i]])
        alone.reset()
        s1 = alone.run(chunk=40)[0]
        o1 = alone.unpack(0)
        alone.close()
        assert s1["steps"] == st[i]["steps"] and s1["failed"] == st[i]["failed"], i
        assert np.array_equal(o1, outs[i]), i


def test_graph_chunks_after_eager_steps_equal_eager():
    """Eager steps on the caller's stream followed at once by graph-replayed
    chunks on batch.run's own stream (which must wait for the caller's
    work) give the same state, bit for bit, as stepping eagerly throughout:
    2D shear layer 512^2, PPM + flattener, 6 + 12 steps."""
    from astrea_amd.plan import Plan
    from astrea_amd import batch, ic
    import torch
    n = 512
    q0 = ic.to_device(ic.shear_layer(n, 20261202), 1.4, "wrap", True)
    prob = dict(n0=n, n1=n, boundary="wrap", dx=1 / n, gamma=1.4, cfl=.4, t_end=1e9, max_steps=18)
    _say(f"[gpu] graph chunks after eager steps, shear {n}^2")
    out = []
    for mixed in (False, True):
        plan = Plan(2, "ppm", "hllc", "ssprk(3,3)", [prob], flattener="astrea", varmask=0x17, compat="physical")
        plan.load([q0])
        plan.reset()
        if mixed:
            plan.step(6)
            st, _, _ = batch.run(plan, chunk=6, use_graph=True)
        else:
            st = plan.run(chunk=18)
        out.append((plan.state.clone(), st[0]["steps"], st[0]["t"]))
        plan.close()
    torch.cuda.synchronize()
    assert out[0][1] == out[1][1] == 18 and out[0][2] == out[1][2]
    assert torch.equal(out[0][0], out[1][0])


@pytest.mark.parametrize("compat", ["astrea", "physical"])
def test_batch_vs_oracle(compat):
    """Every problem of a batch against the oracle, in both compatibility
    modes: a completed problem ends in the oracle's state after the oracle's
    step count; a problem the device flags as failed after n good steps is
    one on which the oracle with the reference's LAPACK eigenvalues raises in
    step n + 1, and the device keeps the oracle's state after n steps."""
    from astrea_amd.plan import Plan
    from astrea_amd import batch
    grids, probs = _ensemble(6, 5, 32, 128)
    plan = Plan(1, "ppm", "hllc", "ssprk(3,3)", probs, varmask=0x13, compat=compat)
    plan.load(grids)
    plan.reset()
    _say(f"[gpu] batch vs oracle, compat={compat}")
    st, _, _ = batch.run(plan, chunk=8)
    outs = plan.unpack()
    plan.close()
    failed = 0
    for i in range(len(grids)):
        p = probs[i]
        s = O.Scheme(dimension=1, subgrid="ppm", solver="hllc", boundary="edge", gamma=1.4, cfl=.5, dx=p["dx"],
                     compat=compat)
        if st[i]["failed"]:
            n_ok = st[i]["steps"]
            _say(f"[gpu]   problem {i} failed after {n_ok} steps")
            lap = replace(s, eig="lapack")
            with pytest.raises(np.linalg.LinAlgError):
                O.run(grids[i], lap, p["t_end"], 1, n_ok + 1)
            ref = O.run(grids[i], s, p["t_end"], 1, n_ok)
            assert ref.steps == n_ok
            assert np.max(rel_l1(outs[i], ref.grid)) <= 1e-12
            failed += 1
            continue
        ref = O.run(grids[i], s, p["t_end"])
        assert st[i]["steps"] == ref.steps
        assert np.max(rel_l1(outs[i], ref.grid)) <= 1e-12
    _say(f"[gpu]   {failed} of {len(grids)} problems failed on the device and in the oracle")


def test_served_results_equal_alone():
    """Continuous batching: requests of two schemes and three sizes admitted
    at different times into re-planned batches give, request by request, the
    same FP64 state (bitwise), step count and failure status as running each
    request alone."""
    from astrea_amd import serve
    reqs = serve.make_requests(8, 3, rate=50.0, sizes=(24, 32, 48))
    for r in reqs:
        r.t_end = .03
    _say(f"[gpu] serve equivalence: {[(r.n, r.subgrid) for r in reqs]}")
    srv = serve.Server(hbm_budget=3 * 48 * 48 * 232, keep_results=True, chunk=4)
    mix = srv.run(reqs, realtime=False)
    res_mix = {r.rid: (r.result.clone(), dict(r.status)) for r in mix["done"]}
    assert mix["rebuilds"] > 2
    alone = serve.run_fifo(reqs, realtime=False, keep_results=True)
    for r in alone["done"]:
        got, st = res_mix[r.rid]
        assert st["steps"] == r.status["steps"] and st["failed"] == r.status["failed"], (r.rid, st, r.status)
        assert torch_equal(got, r.result), r.rid


def test_served_results_equal_alone_at_stream_sizes():
    """The same at the served stream's own sizes (128^2 and 256^2, PLM and
    PPM): each request's state, step count and failure status equal its
    solve alone."""
    from astrea_amd import serve
    reqs = serve.make_requests(6, 5, rate=40.0, sizes=(128, 256))
    for r in reqs:
        r.t_end = .01
    _say(f"[gpu] serve equivalence at stream sizes: {[(r.n, r.subgrid) for r in reqs]}")
    srv = serve.Server(hbm_budget=3 * 256 * 256 * 232, keep_results=True, chunk=4)
    mix = srv.run(reqs, realtime=False)
    res_mix = {r.rid: (r.result.clone(), dict(r.status)) for r in mix["done"]}
    assert mix["rebuilds"] > 2 and len(res_mix) == len(reqs)
    alone = serve.run_fifo(reqs, realtime=False, keep_results=True)
    for r in alone["done"]:
        got, st = res_mix[r.rid]
        assert st["steps"] == r.status["steps"] and st["failed"] == r.status["failed"], (r.rid, st, r.status)
        assert torch_equal(got, r.result), r.rid


def torch_equal(a, b):
    import torch
    return bool(torch.equal(a, b))


def test_quadrants_physical_compat():
    """Seeded four-quadrant requests (PLM and PPM, edge boundaries) with the
    physical HLLC run to t_end on the device and equal the oracle."""
    from astrea_amd import ic
    from astrea_amd.plan import Plan
    import torch
    for sp in ic.quadrant_requests(4, 20261202, sizes=(24,)):
        n = sp["n"]
        q0 = ic.initial_average(torch.from_numpy(ic.quadrants(n, sp["states"])), 1.4, "edge",
                                sp["subgrid"] == "ppm").numpy()
        _say(f"[gpu] quadrants {n}^2 {sp['subgrid']}")
        plan = Plan(2, sp["subgrid"], "hllc", "ssprk(3,3)", [dict(n0=n, n1=n, boundary="edge", dx=1 / n, gamma=1.4,
                                                                    cfl=.4, t_end=.25)], varmask=0x17,
                    compat="physical")
        plan.load([q0])
        plan.reset()
        st = plan.run(64)[0]
        out = plan.unpack(0)
        plan.close()
        s = O.Scheme(dimension=2, subgrid=sp["subgrid"], solver="hllc", boundary="edge", gamma=1.4, cfl=.4, dx=1 / n,
                     compat="physical")
        ref = O.run(q0, s, .25)
        assert not st["failed"] and st["steps"] == ref.steps
        assert np.max(rel_l1(out, ref.grid)) <= 1e-12


@pytest.mark.parametrize("seed", range(10))
def test_random_rectangular_quadrants_vs_oracle(seed):
    """Seeded random cases beyond the configuration table: rectangular grids
    whose sides are no multiple of any tile size, four random quadrant states
    (density 0.1-2, velocities -1..1, pressure 0.05-2; seeds 8 and 9 with
    contrasts up to 100), the subgrid, solver, boundary, compatibility mode
    and flattener cycling with the seed, 8 steps.  The device ends in the
    oracle's state (rel L1 <= 1e-12; 5e-11 with WENO) after the same steps,
    or fails where the oracle with the reference's LAPACK eigenvalues raises
    and keeps the oracle's last state."""
    from astrea_amd.plan import Plan
    rng = np.random.default_rng(7100 + seed)
    n0, n1 = (int(x) for x in rng.integers(17, 42, size=2))
    sg = ["pcm", "plm", "ppm", "weno5", "ppm"][seed % 5]
    solver = ["lf", "hllc"][seed % 2]
    bc = ["edge", "wrap"][(seed // 2) % 2]
    compat = ["astrea", "physical"][(seed // 4) % 2]
    flat = ["off", "astrea", "colella90"][seed % 3] if sg == "ppm" else "off"
    x0, y0 = rng.uniform(.3, .7, size=2)
    x = (np.arange(n0) + .5) / n0
    y = (np.arange(n1) + .5) / n1
    X, Y = np.meshgrid(x, y, indexing="ij")
    quad = (X >= x0).astype(int) + 2 * (Y >= y0).astype(int)
    lo = .01 if seed >= 8 else .1
    states = np.zeros((4, 8))
    states[:, 0] = rng.uniform(lo, 1, 4) * (10 if seed >= 8 else 2)
    states[:, 1:3] = rng.uniform(-1, 1, (4, 2))
    states[:, 4] = rng.uniform(lo / 2, 1, 4) * (10 if seed >= 8 else 2)
    w = states[quad]
    dx = 1 / n0
    s = O.Scheme(dimension=2, subgrid=sg, solver=solver, boundary=bc, gamma=1.4, cfl=.4, dx=dx, compat=compat,
                 flattener=flat)
    q0 = O.initial_average(w, s)
    _say(f"[gpu] random quadrants seed={seed}: {n0}x{n1} {sg}/{solver}/{bc}/{compat}/flattener={flat}")
    prob = dict(n0=n0, n1=n1, boundary=bc, dx=dx, gamma=1.4, cfl=.4, t_end=1e9, max_steps=8)
    plan = Plan(2, sg, solver, "ssprk(3,3)", [prob], varmask=0x17, compat=compat, flattener=flat)
    plan.load([q0])
    plan.reset()
    st = plan.run(16)[0]
    out = plan.unpack(0)
    plan.close()
    tol = 5e-11 if sg == "weno5" else 1e-12
    with np.errstate(all="ignore"):
        if st["failed"]:
            n_ok = st["steps"]
            _say(f"[gpu]   failed after {n_ok} steps")
            with pytest.raises(np.linalg.LinAlgError):
                O.run(q0, replace(s, eig="lapack"), 1e9, 1, n_ok + 1)
            ref = O.run(q0, s, 1e9, 1, n_ok)
        else:
            ref = O.run(q0, s, 1e9, 1, 8)
            assert st["steps"] == ref.steps == 8
    e = float(np.max(rel_l1(out, ref.grid)))
    _say(f"[gpu]   {st['steps']} steps, rel L1 vs oracle {e:.1e}")
    assert e <= tol


@pytest.mark.parametrize("seed", range(6))
def test_random_mhd_2d_vs_oracle(seed):
    """Seeded random 2D MHD fields through the constrained-transport path:
    rectangular grids (12-24 cells a side), Orszag-Tang-like sine fields with
    random amplitudes, phases and a random uniform background, the subgrid
    (PLM/PPM/WENO-5), solver (LLF/HLLC/HLLD), boundary and time stepper
    (SSPRK(3,3)/RK4/SSPRK(2,2)) cycling with the seed, 4 steps: the oracle's
    state (rel L1 <= 1e-12), or a failure where the oracle with the
    reference's LAPACK eigenvalues raises, keeping the oracle's last state."""
    from astrea_amd.plan import Plan
    rng = np.random.default_rng(9700 + seed)
    n0, n1 = (int(x) for x in rng.integers(12, 25, size=2))
    sg = ["plm", "ppm", "weno5"][seed % 3]
    solver = ["lf", "hlld", "hllc"][(seed // 2) % 3]
    bc = ["wrap", "edge"][seed % 2]
    ts = ["ssprk(3,3)", "rk4", "ssprk(2,2)"][(seed // 3) % 3]
    x = (np.arange(n0) + .5) / n0
    y = (np.arange(n1) + .5) / n1
    X, Y = np.meshgrid(x, y, indexing="ij")
    ph = rng.uniform(0, 1, 4)
    a = rng.uniform(.5, 1.5, 4)
    w = np.zeros((n0, n1, 8))
    w[..., 0] = 25 / (36 * np.pi) * (1 + .3 * a[0] * np.sin(2 * np.pi * (X + ph[0])))
    w[..., 4] = 5 / (12 * np.pi) * (1 + .2 * a[1] * np.cos(2 * np.pi * (Y + ph[1])))
    w[..., 1] = -a[2] * np.sin(2 * np.pi * (Y + ph[2])) + rng.uniform(-.3, .3)
    w[..., 2] = a[3] * np.sin(2 * np.pi * (X + ph[3])) + rng.uniform(-.3, .3)
    w[..., 5] = -np.sin(2 * np.pi * (Y + ph[0])) / np.sqrt(4 * np.pi) + rng.uniform(-.1, .1)
    w[..., 6] = np.sin(4 * np.pi * (X + ph[1])) / np.sqrt(4 * np.pi) + rng.uniform(-.1, .1)
    dx = 1 / n0
    s = O.Scheme(dimension=2, subgrid=sg, solver=solver, timestep=ts, boundary=bc, gamma=5 / 3, cfl=.4, dx=dx,
                 magnetic_2d=True)
    q0 = O.initial_average(w, s)
    _say(f"[gpu] random mhd seed={seed}: {n0}x{n1} {sg}/{solver}/{ts}/{bc}")
    prob = dict(n0=n0, n1=n1, boundary=bc, dx=dx, gamma=5 / 3, cfl=.4, t_end=1e9, max_steps=4)
    plan = Plan(2, sg, solver, ts, [prob], varmask=0x77, magnetic_2d=True)
    plan.load([q0])
    plan.reset()
    st = plan.run(8)[0]
    out = plan.unpack(0)
    plan.close()
    with np.errstate(all="ignore"):
        if st["failed"]:
            n_ok = st["steps"]
            _say(f"[gpu]   failed after {n_ok} steps")
            with pytest.raises(np.linalg.LinAlgError):
                O.run(q0, replace(s, eig="lapack"), 1e9, 1, n_ok + 1)
            ref = O.run(q0, s, 1e9, 1, n_ok)
        else:
            ref = O.run(q0, s, 1e9, 1, 4)
            assert st["steps"] == ref.steps == 4
    e = float(np.max(rel_l1(out, ref.grid)))
    _say(f"[gpu]   {st['steps']} steps, rel L1 vs oracle {e:.1e}")
    assert e <= 1e-12


RK_ALL = ["euler", "ssprk(2,2)", "ssprk(3,3)", "ssprk(4,3)", "rk4", "ssprk(5,3)", "ssprk(5,4)", "ssprk(10,4)"]


@pytest.mark.parametrize("seed", range(12))
def test_random_1d_riemann_vs_oracle(seed):
    """Seeded random 1D Riemann problems off the configuration table: the
    time stepper cycles through all eight schemes, the subgrid through
    PCM/PLM/PPM/WENO-3/5/7, with LLF or HLLC, either boundary and
    compatibility mode, 100-700 cells (several tiles), random left/right
    states, 10 steps: the oracle's state after the same steps (rel L1
    <= 1e-12; 5e-11 with WENO-7), or a failure where the oracle with the
    reference's LAPACK eigenvalues raises, keeping the oracle's last state."""
    from astrea_amd.plan import Plan
    rng = np.random.default_rng(9100 + seed)
    n = int(rng.integers(100, 701))
    ts = RK_ALL[seed % 8]
    sg = ["pcm", "plm", "ppm", "weno3", "weno5", "weno7"][seed % 6]
    solver = ["lf", "hllc"][(seed // 2) % 2]
    bc = ["edge", "wrap"][(seed // 3) % 2]
    compat = ["astrea", "physical"][(seed // 5) % 2]
    left, right = np.zeros(8), np.zeros(8)
    for st_ in (left, right):
        st_[0], st_[1], st_[4] = rng.uniform(.1, 2), rng.uniform(-1, 1), rng.uniform(.05, 2)
    x0 = rng.uniform(.3, .7)
    w = np.where(((np.arange(n) + .5) / n < x0)[:, None], left, right)
    s = O.Scheme(dimension=1, subgrid=sg, solver=solver, timestep=ts, boundary=bc, gamma=1.4, cfl=.5, dx=1 / n,
                 compat=compat)
    q0 = O.initial_average(w, s)
    _say(f"[gpu] random 1D seed={seed}: n={n} {sg}/{solver}/{ts}/{bc}/{compat}")
    prob = dict(n0=n, boundary=bc, dx=1 / n, gamma=1.4, cfl=.5, t_end=1e9, max_steps=10)
    plan = Plan(1, sg, solver, ts, [prob], varmask=0x13, compat=compat)
    plan.load([q0])
    plan.reset()
    st = plan.run(16)[0]
    out = plan.unpack(0)
    plan.close()
    tol = 5e-11 if sg == "weno7" else 1e-12
    with np.errstate(all="ignore"):
        if st["failed"]:
            n_ok = st["steps"]
            _say(f"[gpu]   failed after {n_ok} steps")
            with pytest.raises(np.linalg.LinAlgError):
                O.run(q0, replace(s, eig="lapack"), 1e9, 1, n_ok + 1)
            ref = O.run(q0, s, 1e9, 1, n_ok)
        else:
            ref = O.run(q0, s, 1e9, 1, 10)
            assert st["steps"] == ref.steps == 10
    e = float(np.max(rel_l1(out, ref.grid)))
    _say(f"[gpu]   {st['steps']} steps, rel L1 vs oracle {e:.1e}")
    assert e <= tol


MHD = sorted(k for k, v in meta()["cases"].items() if v.get("magnetic_2d"))


@pytest.mark.parametrize("name", MHD)
def test_mhd_constrained_transport_trajectory(name):
    """2D Orszag-Tang with the constrained-transport update (transverse
    reconstruction, corner EMF with the step-parity role flip, EMF-difference
    rows of L, stage-0 face-field reassignment, inverse reconstruction after
    every stage): same steps as the reference, relative L1 <= 1e-12 vs the
    oracle and vs the reference's trajectory."""
    from astrea_amd.plan import Plan
    m, g = meta()["cases"][name], load("steps.npz")
    grid0 = g[name + "_grid0"]
    n = grid0.shape[0]
    _say(f"[gpu] mhd trajectory {name} shape={grid0.shape}")
    plan = Plan(2, m["subgrid"], m["solver"], m["timestep"],
                [dict(n0=n, n1=n, boundary=m["boundary"], dx=m["dx"], gamma=m["gamma"], cfl=m["cfl"],
                      t_end=m["t_end"], max_steps=m["max_steps"])], varmask=0x77, magnetic_2d=True)
    plan.load([grid0])
    plan.reset()
    st = plan.run(chunk=4)[0]
    out = plan.unpack(0)
    plan.close()
    s = O.Scheme(dimension=2, subgrid=m["subgrid"], solver=m["solver"], boundary=m["boundary"], gamma=m["gamma"],
                 cfl=m["cfl"], dx=m["dx"], magnetic_2d=True)
    ref = O.run(grid0, s, m["t_end"], max_steps=m["max_steps"])
    assert st["steps"] == m["steps"] == ref.steps and not st["failed"]
    e_or, e_gold = np.max(rel_l1(out, ref.grid)), np.max(rel_l1(out, g[name + "_final"]))
    _say(f"[gpu]   rel L1 vs oracle {e_or:.2e}, vs reference {e_gold:.2e}")
    assert e_or <= 1e-12 and e_gold <= 1e-12


def test_mhd_longer_run_vs_oracle():
    """Orszag-Tang 32^2, WENO5 + HLLD, 12 steps: device vs oracle."""
    from astrea_amd.plan import Plan
    from astrea_amd import ic
    import torch
    n = 32
    q0 = ic.initial_average(torch.from_numpy(ic.orszag_tang(n)), 5 / 3, "wrap", True).numpy()
    _say("[gpu] mhd OT 32^2 weno5/hlld 12 steps")
    plan = Plan(2, "weno5", "hlld", "ssprk(3,3)", [dict(n0=n, n1=n, boundary="wrap", dx=1 / n, gamma=5 / 3, cfl=.4,
                                                          t_end=10., max_steps=12)], varmask=0x77, magnetic_2d=True)
    plan.load([q0])
    plan.reset()
    st = plan.run(chunk=12)[0]
    out = plan.unpack(0)
    plan.close()
    s = O.Scheme(dimension=2, subgrid="weno5", solver="hlld", boundary="wrap", gamma=5 / 3, cfl=.4, dx=1 / n,
                 magnetic_2d=True)
    ref = O.run(q0, s, 10., max_steps=12)
    assert st["steps"] == 12 and not st["failed"]
    e = np.max(rel_l1(out, ref.grid))
    _say(f"[gpu]   rel L1 vs oracle {e:.2e}")
    assert e <= 1e-12


def test_mhd_orszag_tang_follows_reference_to_failure():
    """Orszag-Tang 64^2, WENO5 + HLLD + constrained transport: the reference
    algorithm reaches a non-finite state at step 17.  The device equals the
    oracle bit for bit through step 16 and raises the failure flag at step
    17, keeping the last finite state (where the reference's LAPACK call
    would raise)."""
    from astrea_amd.plan import Plan
    from astrea_amd import ic
    import torch
    n = 64
    q0 = ic.initial_average(torch.from_numpy(ic.orszag_tang(n)), 5 / 3, "wrap", True).numpy()
    _say("[gpu] mhd OT 64^2 weno5/hlld to failure")
    plan = Plan(2, "weno5", "hlld", "ssprk(3,3)", [dict(n0=n, n1=n, boundary="wrap", dx=1 / n, gamma=5 / 3, cfl=.4,
                                                          t_end=1e9, max_steps=40)], varmask=0x77, magnetic_2d=True,
                compat="physical")
    plan.load([q0])
    plan.reset()
    plan.step(16)
    st16 = plan.status()[0]
    out16 = plan.unpack(0)
    plan.step(24)
    st = plan.status()[0]
    out_last = plan.unpack(0)
    plan.close()
    s = O.Scheme(dimension=2, subgrid="weno5", solver="hlld", timestep="ssprk(3,3)", boundary="wrap", gamma=5 / 3,
                 cfl=.4, dx=1 / n, magnetic_2d=True, compat="physical")
    r16 = O.run(q0, s, 1e9, max_steps=16)
    r17 = O.run(q0, s, 1e9, max_steps=17)
    assert st16["steps"] == 16 and not st16["failed"]
    assert np.array_equal(out16, r16.grid)
    assert not np.isfinite(r17.grid).all()
    assert st["failed"] and st["steps"] == 16
    assert np.array_equal(out_last, out16)


@pytest.mark.parametrize("case", ["sod_ppm", "sod_plm", "quad_plm_edge", "shear_ppm_wrap", "ot_weno5_wrap",
                                  "ot_plm_edge"])
def test_snapshot_equals_reference_conversion(case):
    """ah_plan_snapshot: the primitive snapshot the reference saves every step
    (convert_conservative, transposed by ortho_axis) after a few device steps,
    against the oracle's conversion of the same state, bit for bit."""
    from astrea_amd.plan import Plan
    from astrea_amd import ic
    import torch
    name, sub, *rest = case.split("_")
    bc = rest[0] if rest else "edge"
    _say(f"[gpu] snapshot {case}")
    if name == "sod":
        n, dims, gamma, mask, solver, mhd = 40, 1, 1.4, 0x13, "hllc", False
        w = ic.sod(n)
    elif name == "quad":
        n, dims, gamma, mask, solver, mhd = 19, 2, 1.4, 0x17, "hllc", False
        w = ic.quadrants(n, [(1, .2, .1, 1), (.5, -.3, .2, .4), (.2, .1, -.5, .2), (.8, 0, 0, .9)])
    elif name == "shear":
        n, dims, gamma, mask, solver, mhd = 24, 2, 1.4, 0x17, "hllc", False
        w = ic.shear_layer(n, 3)
    else:
        n, dims, gamma, mask, solver, mhd = 16, 2, 5 / 3, 0x77, "hlld", True
        w = ic.orszag_tang(n)
    ho = sub in ("ppm", "weno5")
    q0 = ic.initial_average(torch.from_numpy(w), gamma, bc, ho).numpy()
    prob = dict(n0=n, boundary=bc, dx=1 / n, gamma=gamma, cfl=.4, t_end=1e9, max_steps=3)
    if dims == 2:
        prob["n1"] = n
    plan = Plan(dims, sub, solver, "ssprk(3,3)", [prob], varmask=mask, magnetic_2d=mhd)
    plan.load([q0])
    plan.reset()
    plan.run(chunk=3)
    snap = plan.snapshot(0).cpu().numpy()
    state = plan.unpack(0)
    plan.close()
    s = O.Scheme(dimension=dims, subgrid=sub, solver=solver, boundary=bc, gamma=gamma, magnetic_2d=mhd)
    ref = O.to_prim(state, s)
    if dims == 2:
        ref = ref.transpose(1, 0, 2)
    assert snap.shape == ref.shape
    assert np.array_equal(snap, ref)


def _series_case(case):
    from astrea_amd import ic, options
    import torch
    name, sub = case.split("_")
    if name == "sod":
        n, t_end, chk, cfl = 64, .05, 4, .5
        w = ic.sod(n)
    else:
        n, t_end, chk, cfl = 16, .1, 3, .4
        w = ic.quadrants(n, [(1, .2, .1, 1), (.5, -.3, .2, .4), (.2, .1, -.5, .2), (.8, 0, 0, .9)])
    sv = options.handle_variables({"config": name, "cells": n, "subgrid": sub, "solver": "hllc",
                                   "timestep": "ssprk(3,3)", "cfl": cfl, "checkpoints": chk}, "edge", t_end=t_end)
    assert sv.dimension == (1 if name == "sod" else 2)
    q0 = ic.initial_average(torch.from_numpy(w), 1.4, "edge", sub == "ppm").numpy()
    s = O.Scheme(dimension=sv.dimension, subgrid=sub, solver="hllc", timestep="ssprk(3,3)", boundary="edge",
                 gamma=1.4, cfl=sv.cfl, dx=sv.dx)
    ref = []
    r = O.run(q0, s, t_end, chk, snapshots=ref)
    return sv, q0, r, ref


@pytest.mark.parametrize("every", ["step", "checkpoint"])
@pytest.mark.parametrize("case", ["sod_plm", "ll3_ppm"])
def test_snapshot_series_equals_reference_loop(case, every, tmp_path):
    """core_run with a SnapshotFile, snapshots taken by the step program:
    every="step" gives one snapshot per pass of the time loop (t = 0, every
    step, t_end) keyed str(float(t)), equal to the oracle's restatement of
    astrea.py:44-50 key for key and bit for bit; every="checkpoint" gives its
    subset at t = 0 and after the steps whose dt was clamped to a checkpoint."""
    from astrea_amd import core
    from astrea_amd.output import SnapshotFile
    _say(f"[gpu] snapshot series {case} every={every}")
    sv, q0, r, ref = _series_case(case)
    path = tmp_path / "series.npz"
    with SnapshotFile(path) as f:
        _, st = core.core_run(q0, sv, save=f, every=every)
    assert st["steps"] == r.steps and not st["failed"]
    if every == "checkpoint":
        assert r.clamped
        ref = [ref[0]] + [ref[k] for k in r.clamped if k < len(ref)]
    with np.load(path) as z:
        keys = sorted(z.files, key=float)
        _say(f"[gpu]   {len(keys)} snapshots, last t={keys[-1]}")
        assert keys == [k for k, _ in ref]
        for k, a in ref:
            assert np.array_equal(z[k], a), k


@pytest.mark.parametrize("name", ["runs.npz", "runs.hdf5"])
def test_save_run_layout(tmp_path, name):
    """core.save_run: file attributes, a group per run with the run
    attributes, the checkpoint series inside it equal to the oracle's -- in
    the reference's HDF5 file (.hdf5) and in the zip archive."""
    from astrea_amd import core, h5
    from astrea_amd.output import read_attrs, read_snapshot, snapshot_keys
    if name.endswith(".hdf5") and not h5.available():
        pytest.skip("libhdf5 cannot be loaded")
    _say(f"[gpu] save_run sod_plm, two runs into one file ({name})")
    sv, q0, r, ref = _series_case("sod_plm")
    path = tmp_path / name
    k1, _, st = core.save_run(path, q0, sv, seed=11, every="checkpoint")
    k2, _, _ = core.save_run(path, q0, sv, seed=11, every="checkpoint")
    assert k1 != k2 and st["steps"] == r.steps
    assert read_attrs(path)["seed"] == 11
    a = read_attrs(path, k1)
    assert (a["config"], a["cells"], a["subgrid"], a["solver"], a["dimension"]) == ("sod", 64, "plm", "hllc", 1)
    assert a["elapsed"] > 0
    want = [ref[0]] + [ref[k] for k in r.clamped if k < len(ref)]
    for key in (k1, k2):
        got = snapshot_keys(path, key)
        assert got == [k for k, _ in want]
        for g, (_, w) in zip(got, want):
            assert np.array_equal(read_snapshot(path, f"{key}/{g}"), w)


@pytest.mark.parametrize("params", [
    {"config": "sod", "cells": 128, "subgrid": "plm", "solver": "hllc", "timestep": "ssprk(3,3)", "cfl": .8},
    {"config": "toro2", "cells": 96, "subgrid": "ppm", "solver": "lf", "timestep": "ssprk(2,2)", "cfl": .5},
    {"config": "ll3", "cells": 24, "dimension": 2, "subgrid": "plm", "solver": "hllc", "timestep": "ssprk(3,3)",
     "cfl": .4, "max_steps": 8},
    {"config": "khi", "cells": 24, "dimension": 2, "subgrid": "ppm", "solver": "hllc", "timestep": "ssprk(3,3)",
     "cfl": .4, "max_steps": 8}])
def test_simulate_configuration(params):
    """core.simulate: the configuration table, the initial state and the
    device loop together, against the oracle's loop on the same initial
    state (same steps, rel L1 <= 1e-12)."""
    from astrea_amd import core, ic
    _say(f"[gpu] simulate {params['config']} {params['subgrid']} {params['solver']}")
    out, st, sv = core.simulate(params)
    s = O.Scheme(dimension=sv.dimension, subgrid=sv.subgrid, solver=sv.solver, timestep=sv.timestep,
                 boundary=sv.boundary, gamma=sv.gamma, cfl=sv.cfl, dx=sv.dx)
    r = O.run(ic.initialise(sv), s, sv.t_end, sv.checkpoints, params.get("max_steps"))
    assert st["steps"] == r.steps and not st["failed"]
    e = np.max(rel_l1(out, r.grid))
    _say(f"[gpu]   {r.steps} steps, t={st['t']:.6g}, rel L1 vs oracle {e:.2e}")
    assert e <= 1e-12


SWEEP_CONFIGS = {"sod": 1, "sedov": 1, "shu-osher": 1, "slow": 1, "sin": 1, "square": 1, "brio-wu": 1,
                 "lin mhd": 1, "orszag-tang": 2, "kelvin-helmholtz": 2, "ll3": 2, "ivc": 2, "mhd rotor": 2}
SWEEP_SUBGRIDS = ["pcm", "plm", "ppm", "weno3", "weno5", "weno7"]
SWEEP_STEPS = ["euler", "rk4", "ssprk(2,2)", "ssprk(3,3)", "ssprk(4,3)", "ssprk(5,3)", "ssprk(5,4)", "ssprk(10,4)"]


def _sweep_cases():
    """Every configuration category x solver the reference can run, each with
    two (subgrid, time stepper) pairs taken in rotation."""
    cases, i = [], 0
    for c, (cfg, dims) in enumerate(SWEEP_CONFIGS.items()):
        for so in ("lf", "lw", "hllc", "hlld", "os", "es"):
            if dims == 2 and (so == "es" or (so == "os" and cfg in ("orszag-tang", "mhd rotor"))):
                continue
            for r in (0, 3):
                cases.append((cfg, dims, so, SWEEP_SUBGRIDS[(i + c + r) % 6], SWEEP_STEPS[(i + r + 1) % 8]))
            i += 1
    return cases


def _floored_l1(a, b, floor=1e-4):
    """Max over variables of the L1 error of a against b, each variable's L1
    norm floored at `floor` times the largest one: a variable that is zero in
    exact arithmetic (v_z, B_z under the Osher-Solomon dissipation) carries
    rounding noise of a different pattern in each implementation."""
    ax = tuple(range(a.ndim - 1))
    num, den = np.abs(a - b).sum(ax), np.abs(b).sum(ax)
    den = np.maximum(den, floor * den.max())
    return float(np.max(num / np.where(den > 0, den, 1.0)))


@pytest.mark.parametrize("case", _sweep_cases(), ids=lambda c: "-".join(map(str, c)).replace(" ", "_"))
def test_option_surface_parity(case):
    """core.simulate over the option surface, 16 cells per axis, 2 steps, from
    each configuration's own initial state: the device equals the oracle run
    with the same routed options (L1 per variable <= 1e-12 of its norm,
    floored; 5e-11 with WENO-7 or Osher-Solomon, whose nonlinear weights and
    SVD carry ulp differences further), or fails at the step where the
    oracle's state first turns non-finite or its LAPACK/SVD call raises,
    keeping the state before it."""
    from astrea_amd import core, ic
    cfg, dims, so, sg, ts = case
    tol = 5e-11 if (sg == "weno7" or so == "os") else 1e-12
    params = dict(config=cfg, cells=16, dimension=dims, solver=so, subgrid=sg, timestep=ts, max_steps=2)
    _say(f"[gpu] option surface {case}")
    out, st, sv = core.simulate(params)
    s = O.Scheme(dimension=sv.dimension, subgrid=sv.subgrid, solver=sv.solver, category=sv.solver_category,
                 timestep=sv.timestep, boundary=sv.boundary, gamma=sv.gamma, cfl=sv.cfl, dx=sv.dx,
                 magnetic_2d=sv.magnetic_2d)
    q0 = ic.initialise(sv)
    # the step the reference raises in: the oracle with LAPACK eigenvalues
    # raises where the reference does (a non-finite Jacobian state, or the
    # Osher-Solomon pseudo-inverse of a non-finite matrix)
    bad = None
    for k in (1, 2):
        try:
            with np.errstate(all="ignore"):
                O.run(q0, replace(s, eig="lapack"), sv.t_end, sv.checkpoints, k)
        except np.linalg.LinAlgError:
            bad = k
            break
    with np.errstate(all="ignore"):
        ref = O.run(q0, s, sv.t_end, sv.checkpoints, 2 if bad is None else bad - 1) if bad != 1 else None
    if bad is not None:
        _say(f"[gpu]   the reference raises in step {bad}; device steps {st['steps']} failed {st['failed']}")
        assert st["failed"] and st["steps"] == bad - 1
        good = q0 if ref is None else ref.grid
        # the state before the failing step, non-finite entries included
        # (a one-stage scheme updates with a non-finite flux and raises in
        # the next step's eigenvalue call)
        fin = np.isfinite(good)
        assert np.array_equal(fin, np.isfinite(out))
        assert _floored_l1(np.where(fin, out, 0), np.where(fin, good, 0)) <= tol
        return
    assert not st["failed"] and st["steps"] == ref.steps
    e = _floored_l1(out, ref.grid)
    _say(f"[gpu]   {ref.steps} steps, L1 vs oracle {e:.1e}")
    assert e <= tol


def test_simulate_fails_where_the_reference_raises():
    """Lax-Liu 12 at 24^2, PPM/HLLC: in step 8 a Jacobian state turns
    non-finite and the reference's LAPACK call raises (the oracle with
    eig="lapack" raises there too); the device run stops failed after 7 steps
    and keeps the state of step 7."""
    from astrea_amd import core, ic
    from dataclasses import replace
    params = {"config": "ll12", "cells": 24, "dimension": 2, "subgrid": "ppm", "solver": "hllc",
              "timestep": "ssprk(3,3)", "cfl": .4, "max_steps": 20}
    _say("[gpu] simulate ll12 ppm hllc to failure")
    out, st, sv = core.simulate(params)
    assert st["failed"] and st["steps"] == 7
    s = O.Scheme(dimension=2, subgrid="ppm", solver="hllc", timestep="ssprk(3,3)", boundary=sv.boundary,
                 gamma=sv.gamma, cfl=sv.cfl, dx=sv.dx, eig="lapack")
    q0 = ic.initialise(sv)
    with pytest.raises(np.linalg.LinAlgError):
        O.run(q0, s, sv.t_end, 1, 8)
    r = O.run(q0, replace(s, eig="closed"), sv.t_end, 1, 7)
    assert np.max(rel_l1(out, r.grid)) <= 1e-12


def test_status_async_matches_status():
    """ah_plan_status_async: statuses queued in stream order between chunks
    of steps, read after the whole run was queued, equal the synchronous ones
    of a second plan read at the same points."""
    from astrea_amd.plan import Plan
    grids, probs = _ensemble(6, 11, 32, 200)
    _say(f"[gpu] status_async, {len(probs)} problems")
    plan = Plan(1, "ppm", "hllc", "ssprk(3,3)", probs, varmask=0x13)
    plan.load(grids)
    plan.reset()
    pend = []
    for _ in range(6):
        plan.step(20)
        pend.append(plan.status_async())
    plan.step(20)
    final = plan.status()
    assert pend[-1].ready()
    assert pend[0].result() != final or all(not s["active"] for s in final)
    again = Plan(1, "ppm", "hllc", "ssprk(3,3)", probs, varmask=0x13)
    again.load(grids)
    again.reset()
    for k in range(6):
        again.step(20)
        assert again.status() == pend[k].result(), k
    plan.close()
    again.close()


def test_snapshot_capture_host_loop_and_graph_agree():
    """The same series three ways on a 2D problem: the host-stepped loop
    (run_saving), the step program's capture with eager chunks, and the
    capture inside replayed step graphs with a ring of 3 slots drained after
    every 2-step replay (slot reuse)."""
    from astrea_amd import core
    from astrea_amd.core import plan_for
    import torch
    sv, q0, r, ref = _series_case("ll3_ppm")
    _say(f"[gpu] capture host loop / eager / graph, {r.steps} steps")
    got = {}
    for how in ("host", "eager", "graph"):
        plan = plan_for(sv, [q0])
        plan.load([q0])
        plan.reset()
        series = []
        save = lambda k, a: series.append((k, a))
        if how == "host":
            core.run_saving(plan, save, sv.t_end)
        elif how == "eager":
            core.run_captured(plan, save, sv.t_end, "step", chunk=4)
        else:
            side = torch.cuda.Stream(plan.device)
            side.wait_stream(torch.cuda.current_stream(plan.device))
            with torch.cuda.stream(side):
                save("0.0", plan.snapshot(0).cpu().numpy())
                plan.capture(0, "step", slots=3)
                while True:
                    plan.step_graph(2)
                    st = plan.status()[0]
                    for t, a in plan.drain():
                        save(str(float(t)), a.cpu().numpy())
                    if not st["active"]:
                        break
                plan.capture(0, None)
            torch.cuda.current_stream(plan.device).wait_stream(side)
        plan.close()
        got[how] = series
    for how, series in got.items():
        assert [k for k, _ in series] == [k for k, _ in ref], how
        for (k, a), (_, b) in zip(series, ref):
            assert np.array_equal(a, b), (how, k)


def test_diagnostics_of_device_snapshots():
    """Total variation, conservation and solution error computed on the
    device from device snapshots of a periodic shear layer, against the
    oracle's restatements of analytic.py on the same state (rounding-level
    agreement: the reductions sum in another order).  The periodic run keeps
    the totals of its state; the snapshot-based conservation diagnostic of the
    reference goes through the 4th-order primitive round trip and drifts by
    O(h^4) except in the density."""
    from astrea_amd.plan import Plan
    from astrea_amd import diagnostics as D, ic
    import torch
    n = 32
    w0 = ic.shear_layer(n, 9)
    q0 = ic.initial_average(torch.from_numpy(w0), 1.4, "wrap", True).numpy()
    _say(f"[gpu] diagnostics of device snapshots, shear {n}^2")
    plan = Plan(2, "ppm", "hllc", "ssprk(3,3)", [dict(n0=n, n1=n, boundary="wrap", dx=1 / n, gamma=1.4, cfl=.4,
                                                       t_end=1e9, max_steps=4)], flattener="astrea", varmask=0x17,
                compat="physical")
    plan.load([q0])
    plan.reset()
    c0 = D.conservation(plan.snapshot(0), 1.4, "wrap", True)
    tot0 = plan.state.sum(dim=1).cpu().numpy()
    plan.run(chunk=4)
    snap = plan.snapshot(0)
    state = plan.unpack(0)
    tot1 = plan.state.sum(dim=1).cpu().numpy()
    plan.close()
    s = O.Scheme(dimension=2, subgrid="ppm", boundary="wrap", gamma=1.4)
    ref = O.to_prim(state, s).transpose(1, 0, 2)
    rel = lambda a, b: np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
    assert rel(D.total_variation(snap, 1.4), O.total_variation(ref, 1.4)) <= 1e-12
    c1 = D.conservation(snap, 1.4, "wrap", True)
    assert rel(c1, O.conservation(ref, s)) <= 1e-12
    theo = w0.transpose(1, 0, 2)
    for p in (0, 2, 11):
        assert rel(D.solution_error(snap, theo, 1.4, p), O.solution_error(ref, theo, 1.4, p)) <= 1e-12
    assert np.all(np.abs(tot1 - tot0) <= 1e-13 * np.maximum(np.abs(tot0), 1))
    assert abs(c1[0] - c0[0]) <= 1e-13 * c0[0]

"""The reference's operator interface on the device, through the C ABI:
evolve_space (with the constrained-transport attachments), compute_L and
evolve_time (num_methods/evolvers.py:12-206), the 1D MHD Riemann solvers,
the PLM slope limiters, the physical-axis HLLC and the option routing of
magnetic configurations, each against the reference's fixtures or the
oracle.  Every case prints its name before touching the GPU."""
import sys
from dataclasses import replace

import numpy as np
import pytest

from oracle import astrea_oracle as O
from tests._golden import load, rel_l1

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _plan(dims, subgrid, solver, timestep, n, bc, gamma, dx, varmask, **kw):
    from astrea_amd.plan import Plan
    prob = dict(n0=n, n1=n if dims == 2 else 1, boundary=bc, dx=dx, gamma=gamma, cfl=.4, t_end=1.)
    return Plan(dims, subgrid, solver, timestep, [prob], varmask=varmask, **kw)


def _scaled_err(a, b):
    scale = np.max(np.abs(b), axis=tuple(range(b.ndim - 1))) + 1e-300
    return float(np.max(np.abs(a - b) / scale))


# ---- 1D MHD Riemann solvers (solvers.py:69-75, 92-138, 142-232) ---------------
MHD_KEYS = [f"mhd_{sg}_{bc}" for sg in ("pcm", "plm", "ppm", "weno3", "weno5", "weno7") for bc in ("edge", "wrap")]


@pytest.mark.parametrize("key", MHD_KEYS)
@pytest.mark.parametrize("solver", ["llf", "hllc", "hlld"])
def test_interface_flux_1d_mhd(key, solver):
    """ah_evolve_space on 1D MHD states (all 8 components): LLF and HLLD
    against the reference's solver outputs (schemes_1d.npz) and the oracle;
    HLLC (what the reference runs for a magnetic configuration given "hllc",
    its solver_category staying "hll") against the oracle; eigmax against
    the reference's LAPACK value."""
    g = load("schemes_1d.npz")
    _, sg, bc = key.split("_")
    grid = g[key + "_grid"]
    N = grid.shape[0]
    _say(f"[gpu] flux1d mhd {key} {solver}")
    plan = _plan(1, sg, solver, "euler", N, bc, 5 / 3, 1 / 32, 0xFF)
    plan.load([grid])
    fl = plan.evolve_space(0)
    plan.close()
    s = O.Scheme(dimension=1, subgrid=sg, solver=solver, boundary=bc, gamma=5 / 3, dx=1 / 32)
    d = O.reconstruct(grid, s, 0)
    lm, _ = O.local_max_eig(d["jac_state"], s, 0)
    F_or = O.riemann(s, 0, lm, None, d)
    got = fl[0]["flux"]
    assert _scaled_err(got, F_or) <= 1e-13
    if solver != "hllc":
        assert _scaled_err(got, g[key + f"_flux_{solver}"]) <= 1e-12
    e_ref = float(g[key + "_eigmax"])
    assert abs(fl[0]["eigmax"] - e_ref) <= 1e-13 * e_ref


# ---- PLM slope limiters (limiters.py:23-49) ----------------------------------
LIMITERS = ["vanleer", "ospre", "vanalbada1", "koren", "superbee"]


@pytest.mark.parametrize("limiter", LIMITERS)
@pytest.mark.parametrize("dims", [1, 2])
def test_slope_limiter_trajectory(limiter, dims):
    """PLM with each non-minmod slope limiter: 1D Sod (edge) and a 2D
    four-quadrant problem (edge), LLF and HLLC, ten steps, equal to the
    oracle (whose limiters are pinned by their textbook phi(r))."""
    from astrea_amd.plan import Plan
    n = 64 if dims == 1 else 20
    c = (np.arange(n) + .5) / n
    if dims == 1:
        w = np.zeros((n, 8))
        w[:] = [.125, 0, 0, 0, .1, 0, 0, 0]
        w[c <= .5] = [1, .3, 0, 0, 1, 0, 0, 0]
        mask = 0x13
    else:
        x, y = np.meshgrid(c, c, indexing="ij")
        w = np.zeros((n, n, 8))
        w[..., 0] = np.where(x < .5, 1., .3) + np.where(y < .5, .4, 0.)
        w[..., 1] = np.where(y < .5, .2, -.3)
        w[..., 2] = np.where(x < .5, -.1, .25)
        w[..., 4] = np.where(x + y < 1, 1., .2)
        mask = 0x17
    for solver in ("llf", "hllc"):
        s = O.Scheme(dimension=dims, subgrid="plm", solver=solver, boundary="edge", gamma=1.4, cfl=.4, dx=1 / n,
                     limiter=limiter)
        q0 = O.initial_average(w, s)
        _say(f"[gpu] plm limiter={limiter} dims={dims} solver={solver}")
        prob = dict(n0=n, n1=n if dims == 2 else 1, boundary="edge", dx=1 / n, gamma=1.4, cfl=.4, t_end=1e9,
                    max_steps=10)
        plan = Plan(dims, "plm", solver, "ssprk(3,3)", [prob], limiter=limiter, varmask=mask)
        plan.load([q0])
        plan.reset()
        st = plan.run(16)[0]
        out = plan.unpack(0)
        plan.close()
        ref = O.run(q0, s, 1e9, max_steps=10)
        assert st["steps"] == ref.steps == 10 and not st["failed"]
        e = np.max(rel_l1(out, ref.grid))
        _say(f"[gpu]   rel L1 vs oracle {e:.1e}")
        assert e <= 1e-12
        # the limiter is really applied: minmod gives a different state
        assert not np.array_equal(out, O.run(q0, replace(s, limiter="minmod"), 1e9, max_steps=10).grid)


# ---- constrained transport: evolve_space attachments (evolvers.py:26-32) --------
@pytest.mark.parametrize("parity", [0, 1])
def test_evolve_space_mhd_2d(parity):
    """2D Orszag-Tang (WENO5, HLLD, constrained transport): interface fluxes,
    eigmax, the corner EMF and the face-averaged field the reference attaches
    (space_2d.npz), at both step parities (the EMF roles flip with it)."""
    g = load("space_2d.npz")
    key = f"weno5_hlld_wrap_p{parity}"
    grid = g[key + "_grid"]
    N = grid.shape[0]
    _say(f"[gpu] evolve_space mhd {key}")
    plan = _plan(2, "weno5", "hlld", "ssprk(3,3)", N, "wrap", 5 / 3, 1 / 12, 0x77, magnetic_2d=True)
    plan.load([grid])
    fl = plan.evolve_space(parity)
    plan.close()
    for ax in (0, 1):
        err = _scaled_err(fl[ax]["flux"], g[key + f"_flux_ax{ax}"])
        _say(f"[gpu]   axis {ax}: flux {err:.1e}")
        assert err <= 1e-12
        e_ref = float(g[key + f"_eigmax_ax{ax}"])
        assert abs(fl[ax]["eigmax"] - e_ref) <= 1e-13 * e_ref
        corner = g[key + f"_corner_ax{ax}"]
        assert np.max(np.abs(fl[ax]["mag_corner"] - corner)) <= 1e-12 * np.max(np.abs(corner))
    fa0 = g[key + "_faceavg_ax0"][..., 5]                 # sweep frame (x, y)
    fa1 = g[key + "_faceavg_ax1"][..., 6].T               # sweep frame (y, x) -> (x, y)
    assert np.max(np.abs(fl[0]["face_avg"] - fa0)) <= 1e-13 * np.max(np.abs(fa0))
    assert np.max(np.abs(fl[1]["face_avg"] - fa1)) <= 1e-13 * np.max(np.abs(fa1))


def _mhd_grid(n):
    c = (np.arange(n) + .5) / n
    x, y = np.meshgrid(c, c, indexing="ij")
    w = np.zeros((n, n, 8))
    w[..., 0] = 25 / (36 * np.pi)
    w[..., 4] = 5 / (12 * np.pi)
    w[..., 1] = -np.sin(2 * np.pi * y)
    w[..., 2] = np.sin(2 * np.pi * x)
    w[..., 5] = -np.sin(2 * np.pi * y) / np.sqrt(4 * np.pi)
    w[..., 6] = np.sin(4 * np.pi * x) / np.sqrt(4 * np.pi)
    return w


# (dims, subgrid, solver, bc, magnetic, timestep)
API_CASES = [
    (1, "plm", "llf", "edge", False, "ssprk(3,3)"),
    (1, "ppm", "hllc", "wrap", False, "rk4"),
    (1, "weno5", "llf", "edge", False, "ssprk(10,4)"),
    (2, "ppm", "hllc", "wrap", False, "ssprk(3,3)"),
    (2, "plm", "llf", "edge", False, "ssprk(5,4)"),
    (2, "weno5", "hlld", "wrap", True, "ssprk(3,3)"),
    (2, "plm", "llf", "edge", True, "rk4"),
    (2, "weno5", "hlld", "wrap", True, "ssprk(10,4)"),
]


def _api_setup(case, parity):
    dims, subgrid, solver, bc, mhd, ts = case
    n = 48 if dims == 1 else 16
    c = (np.arange(n) + .5) / n
    gamma = 5 / 3 if mhd else 1.4
    if mhd:
        w = _mhd_grid(n)
        mask = 0x77
    elif dims == 1:
        w = np.zeros((n, 8))
        w[:] = [.3, -.4, 0, 0, .2, 0, 0, 0]
        w[c <= .5] = [2.5, .7, 0, 0, 4., 0, 0, 0]
        mask = 0x13
    else:
        x, y = np.meshgrid(c, c, indexing="ij")
        w = np.zeros((n, n, 8))
        w[..., 0] = np.where(np.abs(y - .5) < .25, 2., 1.) + .1 * np.sin(2 * np.pi * x)
        w[..., 1] = np.where(np.abs(y - .5) < .25, .5, -.5)
        w[..., 2] = .05 * np.sin(4 * np.pi * x)
        w[..., 4] = 2.5 + .2 * (x < .3)
        mask = 0x17
    s = O.Scheme(dimension=dims, subgrid=subgrid, solver=solver, timestep=ts, boundary=bc, gamma=gamma, cfl=.4,
                 dx=1 / n, magnetic_2d=mhd, parity=parity)
    q0 = O.initial_average(w, replace(s, parity=0))
    plan = _plan(dims, subgrid, solver, ts, n, bc, gamma, 1 / n, mask, magnetic_2d=mhd)
    plan.load([q0])
    return s, q0, plan


@pytest.mark.parametrize("case", API_CASES)
@pytest.mark.parametrize("parity", [0, 1])
def test_compute_L_vs_oracle(case, parity):
    """ah_compute_L (ah_compute_L_ct with constrained transport, whose B rows
    are EMF differences, evolvers.py:51-58) from the device's own fluxes
    against the oracle's compute_L of the oracle's fluxes."""
    dims, subgrid, solver, bc, mhd, ts = case
    if dims == 1 and parity:
        pytest.skip("1D has one sweep order")
    _say(f"[gpu] compute_L {case} parity={parity}")
    s, q0, plan = _api_setup(case, parity)
    fl = plan.evolve_space(parity)
    Ld = plan.compute_L(fl, parity)
    Lg = plan.unpack(0, state=Ld)
    plan.close()
    Lo = O.compute_L(O.evolve_space(q0, s), s)
    err = _scaled_err(Lg, Lo)
    _say(f"[gpu]   max scaled L error {err:.1e}")
    assert err <= 1e-12


def test_compute_L_plain_entry_refuses_constrained_transport():
    """Without the corner EMF the field rows cannot be formed: the plain
    entry point reports that instead of returning wrong rows."""
    from astrea_amd import _lib as L
    s, q0, plan = _api_setup(API_CASES[5], 0)
    fl = plan.evolve_space(0)
    import ctypes
    import torch
    out = torch.zeros_like(plan.state)
    rc = plan.lib.ah_compute_L(plan.h, 0, 0, ctypes.c_void_p(fl[0]["raw"].data_ptr()),
                               ctypes.c_void_p(fl[1]["raw"].data_ptr()), ctypes.c_void_p(out.data_ptr()),
                               plan._stream())
    plan.close()
    assert rc == L.AH_ERR_UNSUPPORTED


@pytest.mark.parametrize("case", API_CASES)
@pytest.mark.parametrize("parity", [0, 1])
def test_evolve_time_vs_oracle(case, parity):
    """ah_evolve_time from the device's evolve_space fluxes and a CFL dt:
    the stage program of the scheme, the in-place field reassignment and
    the refinement after every stage of constrained transport
    (evolvers.py:38-206), against the oracle's evolve_time."""
    dims, subgrid, solver, bc, mhd, ts = case
    if dims == 1 and parity:
        pytest.skip("1D has one sweep order")
    _say(f"[gpu] evolve_time {case} parity={parity}")
    s, q0, plan = _api_setup(case, parity)
    fo = O.evolve_space(q0, s)
    dt = s.cfl * min(s.dx / fo[ax]["eigmax"] for ax in range(dims))
    fl = plan.evolve_space(parity)
    X = plan.state.clone()
    plan.evolve_time(fl, dt, parity, state=X)
    got = plan.unpack(0, state=X)
    # the plan's own state is untouched
    assert np.array_equal(plan.unpack(0), q0)
    plan.close()
    ref = O.evolve_time(np.copy(q0), fo, dt, s)
    e = float(np.max(rel_l1(got, ref)))
    _say(f"[gpu]   rel L1 vs oracle {e:.1e}")
    assert e <= 1e-12


# ---- physical-axis HLLC on the device ----------------------------------------
def _mirror(q):
    return np.ascontiguousarray(q.transpose(1, 0, 2)[..., [0, 2, 1, 3, 4, 5, 6, 7]])


@pytest.mark.parametrize("subgrid,bc,flattener", [("ppm", "wrap", "astrea"), ("plm", "edge", "off")])
def test_physical_hllc_device_run_is_transposition_equivariant(subgrid, bc, flattener):
    """On the device, a physical-axis 2D run from the mirrored initial state
    (axes swapped, v_x <-> v_y) ends in the mirrored final state, bit for
    bit, and both equal the oracle (tile shapes differ between the axes, so
    this also checks that neither sweep depends on its tiling)."""
    from astrea_amd.plan import Plan
    n = 40
    c = (np.arange(n) + .5) / n
    x, y = np.meshgrid(c, c, indexing="ij")
    w = np.zeros((n, n, 8))
    w[..., 0] = 1 + .5 * (x < .4) + .3 * (y > .6)
    w[..., 1] = .3 * np.sin(2 * np.pi * y)
    w[..., 2] = -.2 * np.cos(2 * np.pi * x)
    w[..., 4] = 1 + .5 * (x + y < 1)
    s = O.Scheme(dimension=2, subgrid=subgrid, solver="hllc", boundary=bc, gamma=1.4, cfl=.4, dx=1 / n,
                 compat="physical", flattener=flattener)
    q0 = O.initial_average(w, s)
    outs = []
    for q in (q0, _mirror(q0)):
        _say(f"[gpu] physical hllc equivariance {subgrid} {bc} mirrored={len(outs)}")
        plan = Plan(2, subgrid, "hllc", "ssprk(3,3)", [dict(n0=n, n1=n, boundary=bc, dx=1 / n, gamma=1.4, cfl=.4,
                                                            t_end=1e9, max_steps=8)],
                    varmask=0x17, compat="physical", flattener=flattener)
        plan.load([q])
        plan.reset()
        st = plan.run(8)[0]
        assert st["steps"] == 8 and not st["failed"]
        outs.append(plan.unpack(0))
        plan.close()
    assert np.array_equal(_mirror(outs[0]), outs[1])
    ref = O.run(q0, s, 1e9, max_steps=8)
    assert np.max(rel_l1(outs[0], ref.grid)) <= 1e-12


def test_physical_hllc_device_toro_test_1():
    """Toro's test 1 to t = .2 with the physical-axis HLLC (PPM): the device
    run equals the oracle's, whose L1 density error against the exact
    solution is pinned in the CPU tests."""
    from astrea_amd.plan import Plan
    n = 256
    x = (np.arange(n) + .5) / n
    w = np.zeros((n, 8))
    w[:] = [.125, 0, 0, 0, .1, 0, 0, 0]
    w[x <= .5] = [1, .75, 0, 0, 1, 0, 0, 0]
    s = O.Scheme(dimension=1, subgrid="ppm", solver="hllc", boundary="edge", gamma=1.4, cfl=.5, dx=1 / n,
                 compat="physical")
    q0 = O.initial_average(w, s)
    _say("[gpu] toro test 1, physical hllc, 256 cells")
    plan = Plan(1, "ppm", "hllc", "ssprk(3,3)", [dict(n0=n, boundary="edge", dx=1 / n, gamma=1.4, cfl=.5, t_end=.2)],
                varmask=0x13, compat="physical")
    plan.load([q0])
    plan.reset()
    st = plan.run(64)[0]
    out = plan.unpack(0)
    plan.close()
    ref = O.run(q0, s, .2)
    assert st["steps"] == ref.steps and not st["failed"] and st["t"] == ref.t
    assert np.max(rel_l1(out, ref.grid)) <= 1e-12


# ---- option routing of magnetic configurations -------------------------------
def test_magnetic_config_with_hllc_runs_hllc():
    """Brio-Wu with solver "hllc": the reference resets the name to "lf" but
    keeps solver_category "hll" (generic.py:241, 266-269), and its dispatch
    (solvers.py:15-19) then runs HLLC.  core.simulate routes the same way
    and equals the oracle run with that category."""
    from astrea_amd import core, ic, options
    params = {"config": "brio-wu", "solver": "hllc", "subgrid": "plm", "timestep": "ssprk(3,3)", "cells": 64,
              "cfl": .4, "gamma": 2.0, "max_steps": 12}
    _say("[gpu] brio-wu + hllc through core.simulate")
    out, st, sv = core.simulate(params)
    assert sv.solver == "lf" and sv.solver_category == "hll"
    s = O.Scheme(dimension=1, subgrid="plm", solver="lf", category="hll", timestep="ssprk(3,3)",
                 boundary=sv.boundary, gamma=sv.gamma, cfl=sv.cfl, dx=sv.dx)
    q0 = ic.initialise(sv)
    ref = O.run(q0, s, sv.t_end, max_steps=12)
    assert st["steps"] == ref.steps == 12 and not st["failed"]
    assert np.max(rel_l1(out, ref.grid)) <= 1e-12
    llf = O.run(q0, replace(s, category=""), sv.t_end, max_steps=12)
    assert not np.array_equal(ref.grid, llf.grid)


@pytest.mark.parametrize("config,dims,solver,subgrid", [("sod", 1, "hlld", "plm"), ("sod", 1, "es", "plm"),
                                                         ("kelvin-helmholtz", 2, "hlld", "plm"),
                                                         ("ll3", 2, "hlld", "ppm")])
def test_hydrodynamic_config_with_mhd_solver(config, dims, solver, subgrid):
    """A hydrodynamic configuration given 'hlld' or 'es' runs those solvers in
    the reference on its 8-component arrays (zero field): core.simulate widens
    the device's component set to one they are built for and equals the
    oracle for 8 steps, or fails at the step the reference does (the
    entropy-stable flux turns non-finite on Sod)."""
    from astrea_amd import core, ic
    params = {"config": config, "solver": solver, "subgrid": subgrid, "timestep": "ssprk(3,3)", "cells": 32,
              "dimension": dims, "max_steps": 8}
    _say(f"[gpu] {config} ({dims}D) + {solver} {subgrid} through core.simulate")
    out, st, sv = core.simulate(params)
    s = O.Scheme(dimension=sv.dimension, subgrid=subgrid, solver=solver, timestep="ssprk(3,3)",
                 boundary=sv.boundary, gamma=sv.gamma, cfl=sv.cfl, dx=sv.dx)
    q0 = ic.initialise(sv)
    if solver == "es":
        # the first step whose state the oracle leaves non-finite is the one
        # the reference raises in; the device keeps the state before it
        with np.errstate(all="ignore"):
            bad = next(k for k in range(1, 9) if not np.isfinite(O.run(q0, s, sv.t_end, max_steps=k).grid).all())
        assert st["failed"] and st["steps"] == bad - 1 and np.isfinite(out).all()
        return
    with np.errstate(all="ignore"):
        ref = O.run(q0, s, sv.t_end, max_steps=8)
    assert not st["failed"] and st["steps"] == ref.steps == 8
    e = float(np.max(rel_l1(out, ref.grid)))
    _say(f"[gpu]   rel L1 vs oracle {e:.1e}")
    assert e <= 1e-12


# ---- artificial viscosity (schemes/ppm.py:138-170), opt-in ---------------------
@pytest.mark.parametrize("bc,solver", [("wrap", "hllc"), ("edge", "llf")])
def test_artificial_viscosity_trajectory(bc, solver):
    """1D PPM with the artificial-viscosity flux switched on: 30 steps of a
    compressive profile equal the oracle's (whose viscosity coefficient is
    pinned to the reference's apply_artificial_viscosity), the interface
    fluxes of ah_evolve_space too, and the option changes the result."""
    from astrea_amd.plan import Plan
    n = 96
    x = (np.arange(n) + .5) / n
    w = np.zeros((n, 8))
    w[:, 0] = 1 + .5 * np.exp(-((x - .5) / .1) ** 2)
    w[:, 1] = -np.sin(2 * np.pi * x)
    w[:, 4] = 1. + .2 * (x < .3)
    s = O.Scheme(dimension=1, subgrid="ppm", solver=solver, boundary=bc, gamma=1.4, cfl=.4, dx=1 / n, av=True)
    q0 = O.initial_average(w, s)
    _say(f"[gpu] artificial viscosity {bc} {solver}")
    prob = dict(n0=n, boundary=bc, dx=1 / n, gamma=1.4, cfl=.4, t_end=1e9, max_steps=30)
    plan = Plan(1, "ppm", solver, "ssprk(3,3)", [prob], varmask=0x13, av=True)
    plan.load([q0])
    fl = plan.evolve_space(0)
    d = O.evolve_space(q0, s)
    assert _scaled_err(fl[0]["flux"], d[0]["flux"]) <= 1e-14
    plan.reset()
    st = plan.run(32)[0]
    out = plan.unpack(0)
    plan.close()
    ref = O.run(q0, s, 1e9, max_steps=30)
    assert st["steps"] == ref.steps == 30 and not st["failed"]
    assert np.max(rel_l1(out, ref.grid)) <= 1e-12
    assert not np.array_equal(ref.grid, O.run(q0, replace(s, av=False), 1e9, max_steps=30).grid)


def test_artificial_viscosity_refused_outside_1d_ppm():
    from astrea_amd import _lib as L
    from astrea_amd.plan import Plan
    for dims, sg in ((1, "plm"), (2, "ppm")):
        prob = dict(n0=16, n1=16, boundary="wrap", dx=1 / 16, gamma=1.4, cfl=.4, t_end=1.)
        with pytest.raises(L.AstreaHipError, match="not implemented"):
            Plan(dims, sg, "hllc", "ssprk(3,3)", [prob], varmask=0x13 if dims == 1 else 0x17, av=True)


# ---- PPM authors c / ph (ppm.py:43-71, limiters.py:53-79, 144-201), opt-in ------
@pytest.mark.parametrize("author", ["c", "ph"])
@pytest.mark.parametrize("bc,solver", [("edge", "hllc"), ("wrap", "llf")])
def test_ppm_author_trajectory(author, bc, solver):
    """1D PPM with author 'c' or 'ph' (the grid-wide interface-limiter tests
    recorded by a first pass over the grid each stage): evolve_space fluxes
    and 20 steps of a Riemann problem equal the oracle, whose interface
    states for these authors are pinned to the reference's ppm.run."""
    from astrea_amd.plan import Plan
    n = 300
    x = (np.arange(n) + .5) / n
    w = np.zeros((n, 8))
    w[:] = [.3, -.4, 0, 0, .2, 0, 0, 0]
    w[x <= .5] = [2.5, .7, 0, 0, 4., 0, 0, 0]
    w[:, 0] *= 1 + .1 * np.sin(6 * np.pi * x)
    s = O.Scheme(dimension=1, subgrid="ppm", solver=solver, boundary=bc, gamma=1.4, cfl=.4, dx=1 / n,
                 ppm_author=author)
    q0 = O.initial_average(w, s)
    _say(f"[gpu] ppm author {author} {bc} {solver}")
    prob = dict(n0=n, boundary=bc, dx=1 / n, gamma=1.4, cfl=.4, t_end=1e9, max_steps=20)
    plan = Plan(1, "ppm", solver, "ssprk(3,3)", [prob], varmask=0x13, ppm_author=author)
    plan.load([q0])
    fl = plan.evolve_space(0)
    assert _scaled_err(fl[0]["flux"], O.evolve_space(q0, s)[0]["flux"]) <= 1e-14
    plan.reset()
    st = plan.run(32)[0]
    out = plan.unpack(0)
    plan.close()
    ref = O.run(q0, s, 1e9, max_steps=20)
    assert st["steps"] == ref.steps == 20 and not st["failed"]
    e = float(np.max(rel_l1(out, ref.grid)))
    _say(f"[gpu]   rel L1 vs oracle {e:.1e}")
    assert e <= 1e-12
    assert not np.array_equal(ref.grid, O.run(q0, replace(s, ppm_author="mc"), 1e9, max_steps=20).grid)


@pytest.mark.parametrize("author", ["c", "ph"])
@pytest.mark.parametrize("bc,solver", [("edge", "hllc"), ("wrap", "llf")])
def test_ppm_author_trajectory_2d(author, bc, solver):
    """2D PPM with author 'c' or 'ph' (the grid-wide tests recorded per sweep
    and stage): evolve_space fluxes of both axes and 10 steps of a sheared
    two-state problem equal the oracle."""
    from astrea_amd.plan import Plan
    n = 40
    X, Y = np.meshgrid((np.arange(n) + .5) / n, (np.arange(n) + .5) / n, indexing="ij")
    w = np.zeros((n, n, 8))
    w[..., 0] = np.where((X > .5) ^ (Y > .4), 1.5, .4) * (1 + .1 * np.sin(6 * np.pi * Y))
    w[..., 1] = np.where(Y > .4, .3, -.2)
    w[..., 2] = .1 * np.sin(2 * np.pi * X)
    w[..., 4] = np.where(X > .5, 1., .15)
    s = O.Scheme(dimension=2, subgrid="ppm", solver=solver, boundary=bc, gamma=1.4, cfl=.4, dx=1 / n,
                 ppm_author=author, compat="physical")
    q0 = O.initial_average(w, s)
    _say(f"[gpu] 2D ppm author {author} {bc} {solver}")
    prob = dict(n0=n, n1=n, boundary=bc, dx=1 / n, gamma=1.4, cfl=.4, t_end=1e9, max_steps=10)
    plan = Plan(2, "ppm", solver, "ssprk(3,3)", [prob], varmask=0x17, ppm_author=author, compat="physical")
    plan.load([q0])
    fl = plan.evolve_space(0)
    ref = O.evolve_space(q0, s)
    for ax in (0, 1):
        assert _scaled_err(fl[ax]["flux"], ref[ax]["flux"]) <= 1e-14
    plan.reset()
    st = plan.run(32)[0]
    out = plan.unpack(0)
    plan.close()
    r = O.run(q0, s, 1e9, max_steps=10)
    assert st["steps"] == r.steps == 10 and not st["failed"]
    e = float(np.max(rel_l1(out, r.grid)))
    _say(f"[gpu]   rel L1 vs oracle {e:.1e}")
    assert e <= 1e-12
    assert not np.array_equal(r.grid, O.run(q0, replace(s, ppm_author="mc"), 1e9, max_steps=10).grid)


def test_ppm_author_refused_outside_hydro_fp64():
    from astrea_amd import _lib as L
    from astrea_amd.plan import Plan
    prob = dict(n0=16, n1=16, boundary="wrap", dx=1 / 16, gamma=1.4, cfl=.4, t_end=1.)
    for dims, sg, kw in ((2, "ppm", dict(varmask=0x17, precision="split")), (2, "weno5", dict(varmask=0x17)),
                         (2, "ppm", dict(varmask=0x77, magnetic_2d=True)), (2, "ppm", dict(varmask=0x17, fuse2d=True)),
                         (1, "ppm", dict(varmask=0x13, flattener="astrea"))):
        with pytest.raises(L.AstreaHipError, match="not implemented"):
            Plan(dims, sg, "hllc" if not kw.get("magnetic_2d") else "hlld", "ssprk(3,3)", [prob], ppm_author="ph",
                 **kw)


# ---- tile shapes of the 1D sweep (the autotuner's choices) -----------------------
@pytest.mark.parametrize("subgrid,solver,precision", [("ppm", "hllc", "fp64"), ("plm", "llf", "fp64"),
                                                      ("ppm", "hllc", "split")])
def test_tile_shapes_give_the_same_bits(subgrid, solver, precision):
    """A ragged batch on the 115-cell tiles (128-thread workgroups) ends in
    the same state, step counts and failure flags as on the default 243-cell
    tiles, bit for bit: the tile shape changes the schedule, not the
    arithmetic."""
    import torch
    from astrea_amd import batch, ic
    from astrea_amd.plan import Plan
    specs = ic.ensemble_1d(9, 3, 40, 700)
    probs = [dict(n0=sp["n"], boundary="edge", dx=1 / sp["n"], gamma=1.4, cfl=.5, t_end=sp["t_end"] / 8)
             for sp in specs]
    outs = []
    for tile_s in (0, 115):
        _say(f"[gpu] tile shape {tile_s or 243}: {subgrid}/{solver}/{precision}, sizes {[p['n0'] for p in probs]}")
        plan = Plan(1, subgrid, solver, "ssprk(3,3)", probs, varmask=0x13, precision=precision, tile_s=tile_s)
        for i, sp in enumerate(specs):
            plan.init_riemann(i, sp["left"], sp["right"])
        plan.reset()
        st, _, _ = batch.run(plan, chunk=8)
        outs.append((plan.state.clone(), [(s["steps"], s["failed"]) for s in st]))
        plan.close()
    assert outs[0][1] == outs[1][1]
    assert torch.equal(outs[0][0], outs[1][0])


def test_tile_shape_refused_where_not_built():
    from astrea_amd import _lib as L
    from astrea_amd.plan import Plan
    prob = dict(n0=64, n1=64, boundary="wrap", dx=1 / 64, gamma=1.4, cfl=.4, t_end=1.)
    for kw in (dict(dims=2, subgrid="ppm", varmask=0x17, tile_s=115), dict(dims=1, subgrid="weno5", varmask=0x13,
                                                                           tile_s=115),
               dict(dims=1, subgrid="ppm", varmask=0x13, tile_s=100)):
        dims, sg = kw.pop("dims"), kw.pop("subgrid")
        with pytest.raises(L.AstreaHipError, match="not implemented"):
            Plan(dims, sg, "hllc", "ssprk(3,3)", [prob], **kw)


# ---- one launch per 2D sweep (reconstruction inside the flux tile) --------------
def _grid2d(n0, n1, seed):
    """A sheared two-state layer with a pressure jump across x = 0.5 (the
    flattener and the limiters' extremum branches both act), n0 x n1."""
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid((np.arange(n0) + .5) / n0, (np.arange(n1) + .5) / n1, indexing="ij")
    w = np.zeros((n0, n1, 8))
    w[..., 0] = np.where((X > .5) ^ (Y > .4), 1.5, .4) + .05 * rng.random((n0, n1))
    w[..., 1] = np.where(Y > .4, .3, -.2) + .01 * rng.random((n0, n1))
    w[..., 2] = .1 * np.sin(2 * np.pi * X)
    w[..., 4] = np.where(X > .5, 1., .15) + .01 * rng.random((n0, n1))
    return w


FUSED_CASES = [
    ("ppm", "hllc", "fp64", "astrea", "physical"),
    ("plm", "hllc", "fp64", "off", "physical"),
    ("ppm", "llf", "fp64", "off", "astrea"),
    ("plm", "llf", "split", "off", "astrea"),
    ("ppm", "hllc", "split", "astrea", "physical"),
    ("ppm", "hllc", "fp64", "colella90", "astrea"),
]


@pytest.mark.parametrize("subgrid,solver,precision,flattener,compat", FUSED_CASES)
def test_fused_sweep_gives_the_same_bits(subgrid, solver, precision, flattener, compat):
    """A ragged 2D batch (tile-multiple and ragged sizes, grids smaller than
    one tile, both boundaries) stepped with one launch per sweep ends in the
    same state, step counts, times and failure flags as with the
    reconstruction and flux launches, bit for bit."""
    import torch
    from astrea_amd.plan import Plan
    shapes = [(64, 48, "wrap"), (37, 21, "edge"), (5, 9, "wrap"), (6, 70, "edge"), (33, 17, "wrap")]
    probs = [dict(n0=a, n1=b, boundary=bc, dx=1 / max(a, b), gamma=1.4, cfl=.4, t_end=1., max_steps=9)
             for a, b, bc in shapes]
    grids = []
    for k, (a, b, bc) in enumerate(shapes):
        s = O.Scheme(dimension=2, subgrid=subgrid, boundary=bc, gamma=1.4, dx=1 / max(a, b))
        grids.append(O.initial_average(_grid2d(a, b, 11 + k), s))
    outs = []
    for fuse in (False, True):
        _say(f"[gpu] 2D sweep fused={fuse}: {subgrid}/{solver}/{precision}/{flattener}/{compat}, {shapes}")
        plan = Plan(2, subgrid, solver, "ssprk(3,3)", probs, varmask=0x17, precision=precision,
                    flattener=flattener, compat=compat, fuse2d=fuse)
        plan.load(grids)
        plan.reset()
        st = plan.run(4)
        outs.append((plan.state.clone(), [(s["steps"], s["failed"], s["t"]) for s in st]))
        plan.close()
    assert outs[0][1] == outs[1][1]
    assert all(s[0] == 9 for s in outs[0][1])
    assert torch.equal(outs[0][0], outs[1][0])


@pytest.mark.parametrize("bc,parity", [("wrap", 0), ("edge", 1)])
def test_fused_sweep_interface_fluxes(bc, parity):
    """ah_evolve_space on a one-launch plan: the interface fluxes and eigmax
    of both axes equal the two-launch plan's bit for bit."""
    import torch
    from astrea_amd.plan import Plan
    n0, n1 = 45, 30
    prob = dict(n0=n0, n1=n1, boundary=bc, dx=1 / n0, gamma=1.4, cfl=.4, t_end=1.)
    s = O.Scheme(dimension=2, subgrid="ppm", boundary=bc, gamma=1.4, dx=1 / n0)
    q0 = O.initial_average(_grid2d(n0, n1, 5), s)
    res = []
    for fuse in (False, True):
        _say(f"[gpu] evolve_space fused={fuse} ppm/hllc {bc} {n0}x{n1} parity={parity}")
        plan = Plan(2, "ppm", "hllc", "ssprk(3,3)", [prob], varmask=0x17, flattener="astrea", fuse2d=fuse)
        plan.load([q0])
        res.append(plan.evolve_space(parity))
        plan.close()
    for ax in (0, 1):
        assert torch.equal(res[0][ax]["raw"], res[1][ax]["raw"])
        assert res[0][ax]["eigmax"] == res[1][ax]["eigmax"]


def test_fused_sweep_refused_where_not_built():
    from astrea_amd import _lib as L
    from astrea_amd.plan import Plan
    prob = dict(n0=32, n1=32, boundary="wrap", dx=1 / 32, gamma=1.4, cfl=.4, t_end=1.)
    for dims, sg, sol, vm in ((1, "ppm", "hllc", 0x13), (2, "weno5", "hllc", 0x17), (2, "ppm", "hlld", 0x77),
                              (2, "pcm", "llf", 0x17)):
        p1 = dict(prob, n1=1) if dims == 1 else prob
        with pytest.raises(L.AstreaHipError, match="not implemented"):
            Plan(dims, sg, sol, "ssprk(3,3)", [p1], varmask=vm, fuse2d=True)


# ---- the option surface ---------------------------------------------------------
def test_every_reachable_option_combination_creates_a_plan():
    """Every configuration category (one configuration each, 1D and 2D,
    hydrodynamic and magnetic) x subgrid x solver x time stepper of the
    reference's option surface creates a device plan through core.plan_for,
    except the combinations the reference itself cannot run: Osher-Solomon on
    2D MHD and entropy-stable in 2D, whose reshapes (solvers.py:281,389)
    assume 1D arrays."""
    from astrea_amd import _lib as L, core, ic
    from astrea_amd.options import handle_variables
    configs = {"sod": 1, "sedov": 1, "shu-osher": 1, "slow": 1, "sin": 1, "square": 1, "brio-wu": 1,
               "lin mhd": 1, "orszag-tang": 2, "kelvin-helmholtz": 2, "ll3": 2, "ivc": 2, "mhd rotor": 2}
    steps = ["euler", "rk4", "ssprk(2,2)", "ssprk(3,3)", "ssprk(4,3)", "ssprk(5,3)", "ssprk(5,4)", "ssprk(10,4)"]
    refused = set()
    _say("[gpu] plan creation over the option surface")
    for cfg, dims in configs.items():
        for so in ("lf", "lw", "hllc", "hlld", "os", "es"):
            for sg in ("pcm", "plm", "ppm", "weno3", "weno5", "weno7"):
                for ts in steps:
                    sv = handle_variables({"config": cfg, "cells": 16, "solver": so, "subgrid": sg, "timestep": ts,
                                           "dimension": dims})
                    g = np.asarray(ic.initialise(sv))
                    try:
                        core.plan_for(sv, [g]).close()
                    except L.AstreaHipError:
                        refused.add((cfg, so))
    expect = {(c, "es") for c, d in configs.items() if d == 2} | {("orszag-tang", "os"), ("mhd rotor", "os")}
    assert refused == expect

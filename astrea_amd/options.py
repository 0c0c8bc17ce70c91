"""astrea's option surface for the finite-volume update.

Re-creates the keys of static/.default.yml, the accepted values of
static/.db.json and the validation/derivation rules of
functions/generic.py:159-286 that the update path reads, so a parameters.yml
written for the reference selects the same scheme here.  Keys the build adds:
``limiter`` (PLM slope limiter; the reference hard-codes minmod,
schemes/plm.py:66), ``flattener`` (PPM slope flattener: off | astrea |
colella90), ``precision_mode`` (fp64 | split), ``max_steps`` and ``av``
(artificial viscosity on the 1D PPM fluxes; the reference computes it only
with a flag no option sets and never applies it, so the default False is its
behaviour) and ``ppm_author`` (mc | c | ph, schemes/ppm.py:13; the reference's
evolve_space always passes 'mc').
"""
import itertools
from collections import namedtuple

import numpy as np

from . import _lib as L
from . import configs

# static/.db.json accepted values (the "type" -> "category" -> "accepted" table)
ACCEPTED = {
    "config": {
        "shock": ["sod", "sod shock", "sod-shock", "blast", "sedov", "sedov blast", "sedov wave", "sedov-blast",
                  "sedov-wave"],
        "special": ["shu-osher", "shu osher", "shu", "osher", "so"] +
                   [f"toro{s}{i}" for s in ("", "-", " ") for i in range(1, 6)],
        "interacting": ["slow", "slow-moving", "slow-moving shock", "slow shock", "slow-shock"],
        "smooth": ["sin", "sine", "sine wave", "sine-wave", "gaussian", "gauss", "linear", "lin", "linear mhd",
                   "lin mhd", "linear-mhd", "lin-mhd"],
        "stationary": ["square", "sq", "square-wave", "square wave"],
        "magnetic": ["ryu-jones", "rj", "ryu", "jones", "brio-wu", "bw", "brio", "wu"],
        "magnetic-2D": ["orszag-tang", "orszag", "tang", "ot", "mhd rotor", "mhd-rotor", "rotor", "mhd blast",
                        "mhd-blast", "mhd blast wave", "mhd-blast-wave"],
        "2D": ["kelvin-helmholtz", "kelvin-helmholtz instability", "khi", "vortex", "isentropic-vortex",
               "isentropic vortex", "ivc"] + [f"ll{s}{i}" for s in ("", " ") for i in range(1, 20)]
              + [f"lax-liu{s}{i}" for s in (" ", "") for i in range(1, 20)],
    },
    "dimension": {"": [1, 2]},
    "precision": {"": ["float", "float16", "float32", "float64"]},
    "subgrid": {"pcm": ["pcm", "constant", "c"], "plm": ["plm", "linear", "l"],
                "ppm": ["ppm", "parabolic", "polynomial", "p"],
                "weno": ["weno", "weno3", "weno-3", "weno 3", "weno5", "weno-5", "weno 5", "weno7", "weno-7",
                         "weno 7"]},
    "timestep": {"euler": ["euler", "forward", "forward-euler", "forward euler"],
                 "rk": ["rk4", "runge-kutta", "runge kutta", "runge kutta 4", "runge-kutta 4", "runge-kutta4"],
                 "ssprk": ["ssprk(2,2)", "ssprk22", "ssprk2,2", "ssprk(3,3)", "ssprk33", "ssprk3,3", "ssprk(4,3)",
                           "ssprk43", "ssprk4,3", "ssprk(5,3)", "ssprk53", "ssprk5,3", "ssprk(5,4)", "ssprk54",
                           "ssprk5,4", "ssprk(10,4)", "ssprk104", "ssprk10,4"]},
    "solver": {"lax": ["lf", "friedrich", "lax-friedrich", "llf", "local lax-friedrich", "lw", "lax-wendroff",
                       "wendroff"],
               "hll": ["hllc", "c", "hlld", "d"],
               "complete": ["os", "osher", "solomon", "osher-solomon", "osher solomon", "es", "entropy",
                            "entropy-stable"]},
    "run_type": {"": ["s", "single", "m", "multiple", "multi", "many"]},
    # keys added by this build
    "limiter": {"": ["minmod", "vanleer", "ospre", "vanalbada1", "koren", "superbee"]},
    "flattener": {"": ["off", "astrea", "colella90"]},
    "precision_mode": {"": ["fp64", "split"]},
    "ppm_author": {"": ["mc", "c", "ph"]},
}
# static/.db.json "default" record
DEFAULTS = {"config": "sod", "dimension": 1, "precision": "float64", "subgrid": "pcm", "timestep": "euler",
            "solver": "lf", "run_type": "single", "limiter": "minmod", "flattener": "off", "precision_mode": "fp64",
            "ppm_author": "mc"}

SimVariables = namedtuple("SimVariables", [
    "config", "cells", "cfl", "gamma", "dimension", "precision", "subgrid", "timestep", "solver", "run_type",
    "checkpoints", "limiter", "flattener", "precision_mode", "max_steps", "av", "ppm_author", "config_category",
    "solver_category",
    "magnetic_2d", "permutations", "ortho_axis", "boundary", "dx", "t_end",
    "start_pos", "end_pos", "shock_pos", "misc", "initial_left", "initial_right"])


def _category(key, value):
    for cat, vals in ACCEPTED[key].items():
        if value in vals:
            return cat
    return None


def handle_variables(params, boundary=None, dx=None, t_end=None):
    """generic.py:159-286 for the keys the update reads, merged with the
    configuration's test conditions (static/tests.py, ``configs``) as
    astrea.py:116-119 merges them.  ``params`` is the flattened
    parameters.yml mapping; invalid values revert to the defaults exactly as
    the reference does (with a printed warning).  ``boundary``, ``dx`` and
    ``t_end`` override the configuration's values when given."""
    out = {}
    for k, v in params.items():
        if k in ("checkpoints", "dimension"):
            v = v if isinstance(v, int) and not isinstance(v, bool) else 1
        elif k == "cells":
            try:
                v = int(v) - int(v) % 2
            except (TypeError, ValueError):
                v = 128
        elif k in ("gamma", "cfl"):
            if not isinstance(v, float):
                v = 1.4 if k == "gamma" else .5
        elif k == "av":
            v = v if isinstance(v, bool) else False
        elif k in ACCEPTED:
            v = v.lower() if isinstance(v, str) else v
            if _category(k, v) is None:
                print(f"{k.upper()} value not valid; reverting back to default value: {DEFAULTS[k]}..")
                v = DEFAULTS[k]
        out[k] = v
    for k, v in DEFAULTS.items():
        out.setdefault(k, v)
    out.setdefault("cells", 128)
    out.setdefault("cfl", .5)
    out.setdefault("gamma", 1.4)
    out.setdefault("checkpoints", 1)
    out.setdefault("max_steps", -1)
    out.setdefault("av", False)
    if out["gamma"] == 1:
        # generic.py:193-194: the machine epsilon of the parameters' own precision
        out["gamma"] = out["gamma"] + float(np.finfo(params.get("precision", "float64")).eps)
    if out["precision"] not in ("float", "float64"):
        # constructor.py:24 builds the reference's grid in this dtype; the device
        # update computes in FP64 (or the split mode of precision_mode) whatever it says
        print(f"PRECISION {out['precision']}: the device update runs in float64..")
    cc = _category("config", out["config"])
    out["config_category"] = cc
    out["solver_category"] = _category("solver", out["solver"])
    out["magnetic_2d"] = cc == "magnetic-2D"
    if "2D" in (cc or "") and out["dimension"] != 2:
        out["dimension"] = 2
    if out["solver_category"] == "hll" and out["solver"].endswith("c") and cc in ("magnetic",):
        # generic.py:266-269 resets the solver name only: solver_category stays
        # "hll" (set at generic.py:241), and solvers.py:15-19 dispatches on the
        # category first, so the name "lf" (not ending in "d") still runs HLLC
        print("HLLC solver does not work with magnetic fields present..")
        out["solver"] = DEFAULTS["solver"]
    d = out["dimension"]
    perms = [a for a in itertools.permutations(range(d + 1)) if a[-1] == d]
    out["permutations"] = {i: a for i, a in enumerate(perms)}
    out["ortho_axis"] = perms[-1]
    tc = configs.test_conditions(out["config"], out["cells"], out["gamma"])
    out["boundary"] = boundary if boundary is not None else tc["boundary"]
    out["dx"] = dx if dx is not None else tc["dx"]
    out["t_end"] = t_end if t_end is not None else tc["t_end"]
    for k in ("start_pos", "end_pos", "shock_pos", "misc", "initial_left", "initial_right"):
        out[k] = tc[k]
    return SimVariables(**{k: out[k] for k in SimVariables._fields})


# ---- option strings -> C-ABI enums (evolvers.py:14-21, solvers.py:13-31, evolvers.py:79-206)
def recon_of(subgrid):
    s = subgrid.lower()
    if s.startswith("w"):
        tail = s.replace("-", "").replace(" ", "").split("weno")[-1]
        return {"3": L.AH_WENO3, "7": L.AH_WENO7}.get(tail, L.AH_WENO5)
    if s in ("ppm", "parabolic", "p"):
        return L.AH_PPM
    if s in ("plm", "linear", "l"):
        return L.AH_PLM
    return L.AH_PCM


def solver_of(solver, category=None):
    s = solver.lower()
    category = category or _category("solver", s)
    if category == "hll":
        return L.AH_HLLD if s.endswith("d") else L.AH_HLLC
    if category == "complete":
        # solvers.py:20-24: names starting with "o" select Osher-Solomon, the rest entropy-stable
        return L.AH_DOTS if s.startswith("o") else L.AH_ES
    if s.endswith("w"):
        # solvers.py:27-28.  The device takes the sound speed sqrt(gamma p / rho)
        # as the wave speed (solvers.py:83's comment); the reference's column
        # pick from LAPACK's eigenvalue order is not reproduced.
        return L.AH_LW
    return L.AH_LLF


def rk_of(timestep):
    s = timestep.lower()
    if s.startswith("ssprk"):
        code = s.replace(",", "").replace("(", "").replace(")", "").replace("ssprk", "")
        reg, order = int(code[:-1]), int(code[-1])
        if order == 4:
            return L.AH_SSPRK104 if reg == 10 else L.AH_SSPRK54
        if order == 3:
            return {5: L.AH_SSPRK53, 4: L.AH_SSPRK43}.get(reg, L.AH_SSPRK33)
        return L.AH_SSPRK22
    if s.startswith("r"):
        return L.AH_RK4
    return L.AH_EULER


LIMITERS = {"minmod": L.AH_MINMOD, "vanleer": L.AH_VANLEER, "ospre": L.AH_OSPRE, "vanalbada1": L.AH_VANALBADA1,
            "koren": L.AH_KOREN, "superbee": L.AH_SUPERBEE}
FLATTENERS = {"off": L.AH_FLAT_OFF, "astrea": L.AH_FLAT_ASTREA, "colella90": L.AH_FLAT_COLELLA90}

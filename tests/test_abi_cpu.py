"""CPU checks of the boundary: the in-tree library loads without a GPU and
exports every entry point include/astrea_hip.h declares; the option surface
maps astrea's option strings as the reference routes them."""
import ctypes
import os
import re

import pytest

from astrea_amd import _lib as L
from astrea_amd import options as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "astrea_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(ah_\w+)\s*\(", txt, re.M)))


def test_header_matches_binding_list():
    assert header_functions() == sorted(L.EXPORTS)


def test_library_loads_and_exports_every_symbol():
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("library not built (run python -m astrea_amd._build)")
    lib = ctypes.CDLL(L.LIB_PATH)
    for name in header_functions():
        assert hasattr(lib, name), name
    a, b = ctypes.c_int(), ctypes.c_int()
    assert L.load().ah_version(ctypes.byref(a), ctypes.byref(b)) == 0 and a.value == 1
    assert L.load().ah_status_str(-2).decode().startswith("option combination")


def header_struct_fields(name):
    """Field names of ``typedef struct name {...}`` in the header, in order
    (arrays named without their extent)."""
    txt = open(os.path.join(ROOT, "include", "astrea_hip.h")).read()
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), txt, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        # "long long max_steps", "int n0, n1", "int reserved[1]": the type, then declarators
        first = re.match(r"((?:const\s+)?(?:unsigned\s+)?(?:long\s+long|int|double|float|char)\s*\**)\s*(.*)", decl, re.S)
        names += [re.sub(r"\[\d+\]", "", d).strip(" *") for d in first.group(2).split(",")]
    return names


@pytest.mark.parametrize("cname,py", [("ah_scheme", "AhScheme"), ("ah_problem", "AhProblem"),
                                      ("ah_problem_status", "AhProblemStatus")])
def test_struct_fields_follow_the_header(cname, py):
    """The ctypes mirrors name the header's fields in the header's order."""
    assert [f[0] for f in getattr(L, py)._fields_] == header_struct_fields(cname)


def test_struct_sizes_match_c_layout():
    assert ctypes.sizeof(L.AhScheme) == 16 * 4
    assert ctypes.sizeof(L.AhProblem) == 4 * 4 + 4 * 8 + 8
    assert ctypes.sizeof(L.AhProblemStatus) == 8 * 3 + 2 * 4
    assert ctypes.sizeof(L.AhIc) == 2 * 4 + 2 * 8 + 32 * 8


@pytest.mark.parametrize("s,expect", [("ppm", L.AH_PPM), ("parabolic", L.AH_PPM), ("polynomial", L.AH_PCM),
                                      ("weno", L.AH_WENO5), ("weno-7", L.AH_WENO7), ("weno 3", L.AH_WENO3),
                                      ("l", L.AH_PLM), ("c", L.AH_PCM)])
def test_subgrid_routing(s, expect):
    # evolvers.py:14-21 (the DB accepts 'polynomial' as ppm but the router sends it to PCM)
    assert O.recon_of(s) == expect


@pytest.mark.parametrize("s,expect", [("ssprk(3,3)", L.AH_SSPRK33), ("ssprk104", L.AH_SSPRK104),
                                      ("ssprk5,4", L.AH_SSPRK54), ("ssprk(2,2)", L.AH_SSPRK22),
                                      ("runge-kutta", L.AH_RK4), ("forward euler", L.AH_EULER)])
def test_timestep_routing(s, expect):
    assert O.rk_of(s) == expect


def test_solver_routing():
    assert O.solver_of("hllc") == L.AH_HLLC and O.solver_of("c") == L.AH_HLLC
    assert O.solver_of("d") == L.AH_HLLD
    assert O.solver_of("lax-wendroff") == L.AH_LLF     # solvers.py:28: only names ending in 'w' reach LW
    assert O.solver_of("lw") == L.AH_LW                # names ending in 'w' (solvers.py:27-28)
    # solvers.py:20-24: "complete" names starting with "o" are Osher-Solomon, the rest entropy-stable
    for name in ("os", "osher", "osher-solomon", "osher solomon"):
        assert O.solver_of(name) == L.AH_DOTS
    for name in ("solomon", "es", "entropy", "entropy-stable"):
        assert O.solver_of(name) == L.AH_ES


def test_invalid_values_revert_to_defaults():
    sv = O.handle_variables({"config": "khi", "subgrid": "bogus", "solver": "c", "timestep": "ssprk(3,3)",
                             "cells": 65, "dimension": 1})
    assert sv.subgrid == "pcm" and sv.dimension == 2 and sv.cells == 64
    sv = O.handle_variables({"config": "brio-wu", "solver": "hllc"})
    assert sv.solver == "lf"          # generic.py:266-269
    # ... but the category stays "hll" (generic.py:241), and solvers.py:15-19
    # dispatches on it first: the reference still runs HLLC
    assert sv.solver_category == "hll"
    assert O.solver_of(sv.solver, sv.solver_category) == L.AH_HLLC
    sv = O.handle_variables({"config": "brio-wu", "solver": "hlld"})
    assert sv.solver == "hlld" and O.solver_of(sv.solver, sv.solver_category) == L.AH_HLLD
    sv = O.handle_variables({"config": "sod", "solver": "hllc"})
    assert sv.solver == "hllc" and O.solver_of(sv.solver, sv.solver_category) == L.AH_HLLC


def test_gamma_one_takes_the_epsilon_of_the_stated_precision():
    """generic.py:193-194: gamma == 1 is bumped by np.finfo(precision).eps."""
    import numpy as np
    assert O.handle_variables({"gamma": 1.0, "precision": "float32"}).gamma == 1 + float(np.finfo("float32").eps)
    assert O.handle_variables({"gamma": 1.0}).gamma == 1 + float(np.finfo("float64").eps)
    assert O.handle_variables({"gamma": 1.0, "precision": "float16"}).gamma == 1 + float(np.finfo("float16").eps)
    assert O.handle_variables({"gamma": 1}).gamma == 1.4      # an int is not a valid gamma (generic.py:188-192)


def test_no_kernel_is_instantiated_under_two_fp_modes():
    """A kernel template instantiated in two translation units keeps one
    symbol in the library, and the launch runs whichever copy the loader
    kept.  Units built with different floating-point flags (FP64 units with
    -ffp-contract=off, split units with contraction) must therefore not
    instantiate the same kernel, or an FP64 launch could run a contracted
    copy and lose its bitwise parity."""
    import shutil
    import subprocess
    from astrea_amd import _build
    if not os.path.isdir(_build.OBJ) or shutil.which("nm") is None:
        pytest.skip("objects not built (run python -m astrea_amd._build)")
    where = {}
    for unit, flags in _build.UNITS.items():
        obj = os.path.join(_build.OBJ, unit.replace(".hip", ".o"))
        if not os.path.exists(obj):
            pytest.skip(f"{obj} missing")
        out = subprocess.run(["nm", "--defined-only", obj], capture_output=True, text=True).stdout
        for line in out.splitlines():
            sym = line.split()[-1]
            if sym.startswith("_ZN2ah") and "kernel" in sym:
                where.setdefault(sym, set()).add(tuple(flags))
    mixed = sorted(s for s, f in where.items() if len(f) > 1)
    assert not mixed, mixed[:5]


def test_solver_varmask_widens_for_the_solvers_built_on_more_components():
    from astrea_amd.plan import solver_varmask
    assert solver_varmask(L.AH_VARS_H1, 1, L.AH_HLLD) == L.AH_VARS_ALL
    assert solver_varmask(L.AH_VARS_H1, 1, L.AH_ES) == L.AH_VARS_ALL
    assert solver_varmask(L.AH_VARS_H1, 1, L.AH_HLLC) == L.AH_VARS_H1
    assert solver_varmask(L.AH_VARS_H2, 2, L.AH_HLLD) == L.AH_VARS_M6
    assert solver_varmask(L.AH_VARS_H2, 2, L.AH_DOTS) == L.AH_VARS_H2
    assert solver_varmask(L.AH_VARS_M6, 2, L.AH_HLLD) == L.AH_VARS_M6

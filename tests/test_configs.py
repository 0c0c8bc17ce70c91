"""The configuration table (static/tests.py) and the initial states built
from it (constructor.initialise) against the reference's own outputs
(tests/golden/configs.npz, written by make_golden.py --configs)."""
import numpy as np
import pytest

from astrea_amd import configs, ic, options
from tests._golden import load

CASES = [("sod", 1, 1.4), ("sedov", 1, 1.4), ("shu-osher", 1, 1.4), ("sin", 1, 1.4), ("gauss", 1, 1.4),
         ("lin", 1, 5 / 3), ("lin mhd", 1, 5 / 3), ("slow", 1, 1.4), ("sq", 1, 1.4), ("rj", 1, 5 / 3),
         ("bw", 1, 2.), ("toro1", 1, 1.4), ("toro2", 1, 1.4), ("toro3", 1, 1.4), ("toro4", 1, 1.4),
         ("toro5", 1, 1.4), ("sedov", 2, 1.4), ("gauss", 2, 1.4), ("sod", 2, 1.4), ("khi", 2, 1.4),
         ("ivc", 2, 1.4), ("ot", 2, 5 / 3), ("rotor", 2, 1.4), ("mhd blast", 2, 5 / 3),
                ("isentropic-vortex", 2, 1.4), ("lax-liu 5", 2, 1.4), ("sod shock", 1, 1.4), ("sedov blast", 1, 1.4)] + \
        [(f"ll{k}", 2, 1.4) for k in range(1, 20)]


@pytest.mark.parametrize("config,dim,gamma", CASES)
def test_table_and_initial_state(config, dim, gamma):
    g = load("configs.npz")
    N = 32 if dim == 1 else 12
    key = f"{config.replace(' ', '_')}_{dim}d"
    tc = configs.test_conditions(config, N, gamma)
    for name in ("start_pos", "end_pos", "shock_pos", "t_end", "dx"):
        assert float(tc[name]) == float(g[f"{key}_{name}"]), name
    assert tc["boundary"] == str(g[f"{key}_boundary"])
    assert np.array_equal(tc["initial_left"], g[f"{key}_left"])
    assert np.array_equal(tc["initial_right"], g[f"{key}_right"])
    for subgrid in ("plm", "ppm"):
        sv = options.handle_variables({"config": config, "cells": N, "gamma": gamma, "dimension": dim,
                                       "subgrid": subgrid})
        assert sv.dimension == dim and sv.t_end == tc["t_end"] and sv.boundary == tc["boundary"]
        assert np.array_equal(ic.initialise(sv), g[f"{key}_{subgrid}_grid"]), subgrid


def test_quirks_kept():
    # static/tests.py:255: coeff = -1**index is -1 for Lax-Liu 5 and 6 alike
    assert np.array_equal(configs.test_conditions("ll5", 8, 1.4)["initial_left"],
                          configs.test_conditions("ll6", 8, 1.4)["initial_left"])
    # constructor.py:79-81: the rotor's rotation is written into a copy
    sv = options.handle_variables({"config": "rotor", "cells": 8, "dimension": 2})
    assert not np.any(ic.initial_point(sv)[..., 1:3])

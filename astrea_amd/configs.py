"""The test configurations of astrea (static/tests.py:7-327): domain, shock
position, end time, boundary, the left/right primitive states and the extra
parameters each initial condition reads.

``test_conditions(config, cells, gamma)`` answers what
generate_test_conditions answers, matched by the same substring tests in the
same order (a name can satisfy several; the first match wins), with the same
numbers.  The states are primitive 8-vectors (rho, vx, vy, vz, p, Bx, By, Bz).
"""
import numpy as np


def _st(*w):
    return np.array(w, dtype=float)


# Lax-Liu quadrant problems (static/tests.py:210-305): index -> (top-left and
# top-right states = initial_left / initial_right, bottom_left, bottom_right)
def _lax_liu(index):
    if index in (1, 2):
        top = (_st(.5197, -.7259, 0, 0, .4, 0, 0, 0), _st(1, 0, 0, 0, 1, 0, 0, 0))
        if index == 1:
            return top + (_st(.1072, -.7259, -1.4045, 0, .0439, 0, 0, 0), _st(.2579, 0, -1.4045, 0, .15, 0, 0, 0))
        return top + (_st(1, -.7259, -.7259, 0, 1, 0, 0, 0), _st(.5197, 0, -.7259, 0, .4, 0, 0, 0))
    if index in (5, 6):
        c = -1 ** index   # as written: -(1 ** index) = -1 for both
        return (_st(2, c * .75, .5, 0, 1, 0, 0, 0), _st(1, c * .75, -.5, 0, 1, 0, 0, 0),
                _st(1, -c * .75, .5, 0, 1, 0, 0, 0), _st(3, -c * .75, -.5, 0, 1, 0, 0, 0))
    if index in (17, 18, 19):
        v1, v4 = {17: (-.4, -1.1259), 18: (1, .2741)}.get(index, (.3, -.4259))
        return (_st(2, 0, -.3, 0, 1, 0, 0, 0), _st(1, 0, v1, 0, 1, 0, 0, 0),
                _st(1.0625, 0, .2145, 0, .4, 0, 0, 0), _st(.5197, 0, v4, 0, .4, 0, 0, 0))
    table = {
        3: ((1.5, 0, 0, 1.5), (.5323, 1.206, 0, .3), (.5323, 0, 1.206, .3), (.138, 1.206, 1.206, .029)),
        4: ((.5065, .8939, 0, .35), (1.1, 0, 0, 1.1), (1.1, .8939, .8939, 1.1), (.5065, 0, .8939, .35)),
        7: ((.5197, -.6259, .1, .4), (1, .1, .1, 1), (.8, .1, .1, .4), (.5197, .1, -.6259, .4)),
        8: ((1, -.6259, .1, 1), (.5197, .1, .1, .4), (.8, .1, .1, 1), (1, .1, -.6259, 1)),
        9: ((2, 0, -.3, 1), (1, 0, .3, 1), (1.039, 0, -.8133, .4), (.5197, 0, -.4259, .4)),
        10: ((.5, 0, .6076, 1), (1, 0, .4297, 1), (.2281, 0, -.6076, .3333), (.4562, 0, -.4297, .3333)),
        11: ((.5313, .8276, 0, .4), (1, .1, 0, 1), (.8, .1, 0, .4), (.5313, .1, .7276, .4)),
        12: ((1, .7276, 0, 1), (.5313, 0, 0, .4), (.8, 0, 0, 1), (1, 0, .7276, 1)),
        13: ((2, .3, 0, 1), (1, 0, -.3, 1), (1.0625, 0, .8145, .4), (.5313, 0, .4276, .4)),
        14: ((1, 0, -1.2172, 8), (2, 0, -.5606, 8), (.4736, 0, 1.2172, 2.6667), (.9474, 0, 1.1606, 2.6667)),
        15: ((.5197, -.6259, -.3, .4), (1, .1, -.3, 1), (.8, .1, -.3, .4), (.5313, .1, .4276, .4)),
        16: ((1.0222, -.6179, .1, 1), (.5313, .1, .1, .4), (.8, .1, .1, 1), (1, .1, .8276, 1)),
    }
    if index not in table:
        raise ValueError(f"Lax-Liu configuration {index} does not exist (1-19)")
    return tuple(_st(r, u, v, 0, p, 0, 0, 0) for r, u, v, p in table[index])


def test_conditions(config, cells, gamma):
    """static/tests.py:7-327 for `config` (a lower-case name the option table
    accepts), `cells` and `gamma`."""
    misc = None
    if "sod" in config:
        lo, hi, x0, t_end, bc = 0, 1, .5, .2, "edge"
        left, right = _st(1, 0, 0, 0, 1, 0, 0, 0), _st(.125, 0, 0, 0, .1, 0, 0, 0)
    elif "sedov" in config or config == "blast":
        lo, hi, x0, t_end, bc = -10, 10, .5, .6, "wrap"
        left, right = _st(1, 0, 0, 0, 100, 0, 0, 0), _st(1, 0, 0, 0, 1, 0, 0, 0)
    elif "shu" in config or "osher" in config or config == "so":
        lo, hi, x0, t_end, bc = -1, 1, -.8, .47, "edge"
        left, right = _st(3.857143, 2.629369, 0, 0, 10.3333, 0, 0, 0), _st(0, 0, 0, 0, 1, 0, 0, 0)
        misc = {"freq": 5, "ampl": .2, "y_offset": 1}
    elif config.startswith("sin"):
        lo, hi, x0, t_end, bc = 0, 1, 1, 1, "wrap"
        left = right = _st(0, 1, 1, 1, 1, 0, 0, 0)
        misc = {"freq": 2, "ampl": .1, "y_offset": 2}
    elif config.startswith("gauss"):
        lo, hi, x0, t_end, bc = -1, 1, 1, 2, "wrap"
        left = right = _st(0, 1, 1, 1, 1e-6, 0, 0, 0)
        misc = {"peak_pos": 0, "ampl": .75, "fwhm": .08, "y_offset": 1}
    elif config.startswith("lin"):
        lo, hi, x0, t_end, bc = 0, 1, 1, 2 * np.pi, "wrap"
        left = _st(1, 1, 1, 1, 1 / gamma, 0, 0, 0)
        if "mhd" in config:
            left[5:] = np.array([1, np.sqrt(2), .5]) * np.sqrt(4 * np.pi)
        right = left.copy()
        misc = {"freq": 2, "ampl": 1e-6}
    elif "slow" in config:
        lo, hi, x0, t_end, bc = 0, 1, .5, .08, "edge"
        left, right = _st(5.6698, -1.5336, 0, 0, 100, 0, 0, 0), _st(1, -10.5636, 0, 0, 1, 0, 0, 0)
    elif config.startswith("sq"):
        lo, hi, x0, t_end, bc = -1, 1, 1 / 3, .05, "wrap"
        left, right = _st(1, 1, 0, 0, 1, 0, 0, 0), _st(.01, 1, 0, 0, 1, 0, 0, 0)
    elif "ryu" in config or "jones" in config or config == "rj":
        lo, hi, x0, t_end, bc = -.5, .5, 0, .15, "edge"
        s = 1 / np.sqrt(np.pi)
        left = _st(1.08, 1.2, .01, .5, .95, s, 1.8 / np.sqrt(np.pi), s)
        right = _st(1, 0, 0, 0, 1, s, 2 / np.sqrt(np.pi), s)
    elif "brio" in config or "wu" in config or config == "bw":
        lo, hi, x0, t_end, bc = -.5, .5, 0, .1, "edge"
        left, right = _st(1, 0, 0, 0, 1, .75, 1, 0), _st(.125, 0, 0, 0, .1, .75, -1, 0)
    elif "kelvin" in config or "helmholtz" in config or config == "khi":
        lo, hi, x0, t_end, bc = -1, 1, 0, 4, "wrap"
        left, right = _st(2, -.5, 0, 0, 1, 0, 0, 0), _st(1, .5, 0, 0, 1, 0, 0, 0)
        misc = {"perturb_ampl": .5, "freq": 4}
    elif "isentropic" in config or "vortex" in config or config == "ivc":
        lo, hi, x0, t_end, bc = 0, 10, 5, 1, "wrap"
        left = right = _st(1, 0, 0, 0, 1, 0, 0, 0)
        misc = {"vortex_str": 5, "freq": 2}
    elif "orszag" in config or "tang" in config or config == "ot":
        lo, hi, x0, t_end, bc = -.5, .5, .5, np.pi / 2, "wrap"
        left = right = _st(25 / (36 * np.pi), 0, 0, 0, 5 / (12 * np.pi), 0, 0, 0)
        misc = {"ampl": 1 / np.sqrt(4 * np.pi)}
    elif "rotor" in config:
        lo, hi, x0, t_end, bc = -.5, .5, .1, .15, "wrap"
        left, right = _st(10, 0, 0, 0, 1, 0, 0, 0), _st(1, 0, 0, 0, 1, 0, 0, 0)
        misc = {"omega": 20}
    elif "blast" in config and config.startswith("mhd"):
        lo, hi, x0, t_end, bc = -.5, .5, .1, .2, "wrap"
        b = 1 / np.sqrt(2)
        left, right = _st(1, 0, 0, 0, 10, b, b, 0), _st(1, 0, 0, 0, .1, b, b, 0)
    elif "toro" in config:
        lo, hi, bc = 0, 1, "edge"
        if "2" in config:
            x0, t_end = .5, .14
            left, right = _st(1, -2, 0, 0, .4, 0, 0, 0), _st(1, 2, 0, 0, .4, 0, 0, 0)
        elif "3" in config:
            x0, t_end = .5, .012
            left, right = _st(1, 0, 0, 0, 1000, 0, 0, 0), _st(1, 0, 0, 0, .01, 0, 0, 0)
        elif "4" in config:
            x0, t_end = .3, .05
            left, right = _st(5.99924, 19.5975, 0, 0, 460.894, 0, 0, 0), _st(5.99242, -6.19633, 0, 0, 46.095, 0, 0, 0)
        elif "5" in config:
            x0, t_end = .8, .012
            left, right = _st(1, -19.59745, 0, 0, 1000, 0, 0, 0), _st(1, -19.59745, 0, 0, .01, 0, 0, 0)
        else:
            x0, t_end = .3, .2
            left, right = _st(1, .75, 0, 0, 1, 0, 0, 0), _st(.125, 0, 0, 0, .1, 0, 0, 0)
    elif "lax" in config or "liu" in config or "ll" in config:
        lo, hi, x0, t_end, bc = 0, 1, .5, 2, "wrap"
        c = config.replace(" ", "")
        index = int(c.split("ll")[-1]) if "ll" in config else int(c.split("liu")[-1])
        left, right, bl, br = _lax_liu(index)
        misc = {"bottom_left": bl, "bottom_right": br}
    else:
        lo, hi, x0, t_end, bc = 0, 1, .5, .2, "edge"
        left, right = _st(1, 0, 0, 0, 1, 0, 0, 0), _st(.125, 0, 0, 0, .1, 0, 0, 0)
    return {"start_pos": lo, "end_pos": hi, "shock_pos": x0, "t_end": t_end, "boundary": bc, "misc": misc,
            "initial_left": left, "initial_right": right, "dx": abs(hi - lo) / cells, "dy": abs(hi - lo) / cells}

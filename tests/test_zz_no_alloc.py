"""ah_evolve_space_static, the entry point that allocates nothing, against
ah_evolve_space on a plan of the same problem: the same fluxes and eigmax
bit for bit, on two calls in a row whose second problem is smaller (the
static array keeps the first call's contents), in 1D and 2D; a problem
beyond the array's size is refused."""
import ctypes
import sys

import numpy as np
import pytest

from oracle import astrea_oracle as O

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(dims, n, seed):
    rng = np.random.default_rng(seed)
    c = (np.arange(n) + .5) / n
    if dims == 1:
        w = np.zeros((n, 8))
        w[:] = [.3, -.4, 0, 0, .2, 0, 0, 0]
        w[c <= .5] = [2.5, .7, 0, 0, 4., 0, 0, 0]
        w[:, 0] *= 1 + .05 * rng.random(n)
        return w
    x, y = np.meshgrid(c, c, indexing="ij")
    w = np.zeros((n, n, 8))
    w[..., 0] = np.where(np.abs(y - .5) < .25, 2., 1.) + .05 * rng.random((n, n))
    w[..., 1] = np.where(np.abs(y - .5) < .25, .5, -.5)
    w[..., 2] = .01 * np.sin(4 * np.pi * x) + 1e-3 * rng.random((n, n))
    w[..., 4] = 2.5
    return w


def _static_call(plan, parity, fuse2d=False):
    """ah_evolve_space_static with the plan's scheme (or its one-launch 2D
    form) and problem on the plan's state, into fresh output tensors."""
    import torch
    from astrea_amd import _lib as L
    sch = L.AhScheme.from_buffer_copy(plan.scheme)
    sch.fuse2d = 1 if fuse2d else 0
    p = plan.problems[0]
    f0 = torch.full((plan.nv, (p.n0 + 1) * p.n1), np.nan, dtype=torch.float64, device=plan.device)
    f1 = torch.full((plan.nv, (p.n1 + 1) * p.n0), np.nan, dtype=torch.float64, device=plan.device)
    eig = (ctypes.c_double * 2)()
    rc = plan.lib.ah_evolve_space_static(ctypes.byref(sch), ctypes.byref(p), int(parity),
                                         ctypes.c_void_p(plan.state.data_ptr()), ctypes.c_void_p(f0.data_ptr()),
                                         ctypes.c_void_p(f1.data_ptr()), eig, plan._stream())
    return rc, f0, f1, [eig[0], eig[1]]


@pytest.mark.parametrize("dims,subgrid,solver,bc,sizes,flattener,fuse2d,author", [
    (1, "ppm", "hllc", "edge", (4000, 700), "off", False, "mc"),
    (2, "ppm", "hllc", "wrap", (160, 72), "astrea", False, "mc"),
    (2, "plm", "llf", "edge", (97, 41), "off", False, "mc"),
    (2, "ppm", "hllc", "edge", (120, 53), "astrea", True, "mc"),
    (2, "ppm", "llf", "wrap", (64, 40), "off", False, "ph"),
])
def test_static_evolve_space_equals_plan_entry(dims, subgrid, solver, bc, sizes, flattener, fuse2d, author):
    from astrea_amd import _lib as L
    from astrea_amd.plan import Plan
    for k, n in enumerate(sizes):
        for parity in ((0, 1) if dims == 2 else (0,)):
            _say(f"[gpu] static evolve_space dims={dims} {subgrid}/{solver}/{bc} n={n} call={k} parity={parity}"
                 f" fused={fuse2d} author={author}")
            s = O.Scheme(dimension=dims, subgrid=subgrid, boundary=bc, gamma=1.4, dx=1 / n)
            q0 = O.initial_average(_case(dims, n, 7 + k), s)
            prob = dict(n0=n, n1=n if dims == 2 else 1, boundary=bc, dx=1 / n, gamma=1.4, cfl=.4, t_end=1.)
            plan = Plan(dims, subgrid, solver, "ssprk(3,3)", [prob], varmask=0x13 if dims == 1 else 0x17,
                        compat="physical", flattener=flattener, ppm_author=author)
            plan.load([q0])
            ref = plan.evolve_space(parity)
            rc, f0, f1, eig = _static_call(plan, parity, fuse2d)
            assert rc == L.AH_OK, L.load().ah_status_str(rc)
            import torch
            assert torch.equal(f0, ref[0]["raw"])
            assert eig[0] == ref[0]["eigmax"]
            if dims == 2:
                assert torch.equal(f1, ref[1]["raw"])
                assert eig[1] == ref[1]["eigmax"]
            plan.close()


def test_static_evolve_space_refuses_a_problem_beyond_the_array():
    from astrea_amd import _lib as L
    from astrea_amd.plan import Plan
    n = 1024     # 2D FP64 hydrodynamics: 136 B per cell, 136 MiB > 64 MiB
    prob = dict(n0=n, n1=n, boundary="wrap", dx=1 / n, gamma=1.4, cfl=.4, t_end=1.)
    _say("[gpu] static evolve_space refuses 1024^2")
    plan = Plan(2, "ppm", "hllc", "ssprk(3,3)", [prob], varmask=0x17)
    rc, _, _, _ = _static_call(plan, 0)
    plan.close()
    assert rc == L.AH_ERR_TOO_LARGE
    assert "static scratch" in L.load().ah_status_str(rc).decode()

"""Drop-in for astrea's time loop on the device.

``core_run(grid, sim_variables)`` replaces astrea.py:31-85 (without its HDF5
snapshot and plotting I/O): the initial conservative grid (the output of
constructor.initialise) is advanced to ``t_end`` (or ``max_steps``) by the
device plan, which keeps t, dt and the termination test on the GPU.
"""
from .plan import Plan, infer_varmask


def plan_for(sim_variables, grids, precision=None):
    sv = sim_variables
    dims = sv.dimension
    problems = []
    for g in grids:
        problems.append(dict(n0=g.shape[0], n1=g.shape[1] if dims == 2 else 1, boundary=sv.boundary, dx=sv.dx,
                             gamma=sv.gamma, cfl=sv.cfl, t_end=sv.t_end, checkpoints=sv.checkpoints,
                             max_steps=sv.max_steps))
    mask = infer_varmask(grids, dims, sv.magnetic_2d)
    return Plan(dims, sv.subgrid, sv.solver, sv.timestep, problems, limiter=sv.limiter, flattener=sv.flattener,
                precision=precision or sv.precision_mode, varmask=mask, magnetic_2d=sv.magnetic_2d)


def core_run(grid, sim_variables, precision=None, chunk=64):
    """Returns (final grid in the reference layout, status dict)."""
    plan = plan_for(sim_variables, [grid], precision)
    plan.load([grid])
    plan.reset()
    st = plan.run(chunk)
    out = plan.unpack(0)
    plan.close()
    return out, st[0]

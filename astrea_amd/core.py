"""Drop-in for astrea's time loop on the device.

``core_run(grid, sim_variables)`` replaces astrea.py:31-85 (without its
plotting): the initial conservative grid (the output of
constructor.initialise) is advanced to ``t_end`` (or ``max_steps``) by the
device plan, which keeps t, dt and the termination test on the GPU.  With a
``save`` sink it also hands over the snapshot of every pass of the loop, as
the reference writes one HDF5 dataset per pass (astrea.py:46-50).
"""
import torch

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


def run_saving(plan, save, t_end, i=0):
    """Step problem ``i`` of ``plan`` one step per pass and call
    ``save(str(float(t)), snapshot)`` on every pass with t <= t_end, before
    the stop test (astrea.py:44-50): the initial state, each step's state, and
    the state at t_end when the run ends there exactly.  A step that fails
    (non-finite state) ends the loop without a snapshot, where the reference
    raises.  The snapshot is ``Plan.snapshot`` copied to pinned host memory."""
    dev = plan.snapshot(i)
    host = torch.empty(tuple(dev.shape), dtype=torch.float64, pin_memory=True)
    while True:
        st = plan.status()[i]
        if st["failed"]:
            return st
        if st["t"] <= t_end:
            plan.snapshot(i, out=dev)
            host.copy_(dev)
            save(str(float(st["t"])), host.numpy().copy())
        if not st["active"]:
            return st
        plan.step(1)


def core_run(grid, sim_variables, precision=None, chunk=64, save=None):
    """Returns (final grid in the reference layout, status dict).  ``save``
    (e.g. ``output.SnapshotFile``) receives (str(float(t)), snapshot) on every
    pass of the loop; without it the plan steps ``chunk`` steps between looks
    at the device time state."""
    plan = plan_for(sim_variables, [grid], precision)
    plan.load([grid])
    plan.reset()
    st = plan.run(chunk)[0] if save is None else run_saving(plan, save, sim_variables.t_end)
    out = plan.unpack(0)
    plan.close()
    return out, st

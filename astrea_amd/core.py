"""Drop-in for astrea's time loop on the device.

``core_run(grid, sim_variables)`` replaces astrea.py:31-85 (without its
plotting): the initial conservative grid (the output of
constructor.initialise) is advanced to ``t_end`` (or ``max_steps``) by the
device plan, which keeps t, dt and the termination test on the GPU.  With a
``save`` sink it also hands over the snapshot of every pass of the loop, as
the reference writes one HDF5 dataset per pass (astrea.py:46-50).
"""
import os
from datetime import datetime
from time import perf_counter

import torch

from . import ic, options
from .output import open_snapshots
from .plan import Plan, infer_varmask, solver_varmask


def plan_for(sim_variables, grids, precision=None):
    sv = sim_variables
    dims = sv.dimension
    problems = []
    for g in grids:
        problems.append(dict(n0=g.shape[0], n1=g.shape[1] if dims == 2 else 1, boundary=sv.boundary, dx=sv.dx,
                             gamma=sv.gamma, cfl=sv.cfl, t_end=sv.t_end, checkpoints=sv.checkpoints,
                             max_steps=sv.max_steps))
    mask = solver_varmask(infer_varmask(grids, dims, sv.magnetic_2d), dims,
                          options.solver_of(sv.solver, sv.solver_category))
    return Plan(dims, sv.subgrid, sv.solver, sv.timestep, problems, limiter=sv.limiter, flattener=sv.flattener,
                precision=precision or sv.precision_mode, varmask=mask, magnetic_2d=sv.magnetic_2d,
                solver_category=sv.solver_category, av=sv.av, ppm_author=sv.ppm_author)


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


def run_captured(plan, save, t_end, every="step", chunk=16, i=0):
    """The series of ``run_saving`` (every="step") or its subset at t = 0 and
    the checkpoints (every="checkpoint"), with the snapshots taken by the
    step program itself (``Plan.capture``): the host looks at the device once
    per chunk of steps, copies the chunk's snapshots to pinned memory ahead of
    the next chunk on the stream, and calls ``save`` while that chunk runs."""
    st = plan.status()[i]
    if st["t"] <= t_end and not st["failed"]:
        save(str(float(st["t"])), plan.snapshot(i).cpu().numpy())
    plan.capture(i, every, slots=chunk)
    host = torch.empty((chunk,) + plan.shape(i)[::-1] + (8,), dtype=torch.float64, pin_memory=True)
    stream = torch.cuda.current_stream(plan.device)
    try:
        plan.step(chunk)
        while True:
            st = plan.status()[i]
            got = plan.drain()
            for k, (_, snap) in enumerate(got):
                host[k].copy_(snap, non_blocking=True)
            copied = torch.cuda.Event()
            copied.record(stream)
            if st["active"]:
                plan.step(chunk)
            copied.synchronize()
            for k, (t, _) in enumerate(got):
                save(str(float(t)), host[k].numpy().copy())
            if not st["active"]:
                return st
    finally:
        plan.capture(i, None)


def core_run(grid, sim_variables, precision=None, chunk=64, save=None, every="step"):
    """Returns (final grid in the reference layout, status dict).  ``save``
    (e.g. ``output.SnapshotFile``) receives (str(float(t)), snapshot) for
    t = 0 and every step (every="step": the reference's series, astrea.py:46-50)
    or every checkpoint (every="checkpoint"); the step program takes the
    snapshots (``run_captured``).  Without ``save`` the plan steps ``chunk``
    steps between looks at the device time state."""
    plan = plan_for(sim_variables, [grid], precision)
    plan.load([grid])
    plan.reset()
    if save is None:
        st = plan.run(chunk)[0]
    else:
        st = run_captured(plan, save, sim_variables.t_end, every, min(chunk, 16))
    out = plan.unpack(0)
    plan.close()
    return out, st


def save_run(path, grid, sim_variables, seed=None, every="step", precision=None):
    """One run of astrea.py:150-186 into a snapshot file: the file attributes
    datetime ('%Y%m%d%H%M' of the first run) and seed, a group named by the
    run's access key ('%Y%m%d%H%M%S' + microseconds) holding the run's
    snapshot series and the attributes config, cells, cfl, gamma, dimension,
    subgrid, timestep, solver and elapsed (seconds of core_run).  Appends to
    an existing file.  A .hdf5 / .h5 path is the reference's HDF5 file
    (written through libhdf5), any other name the zip archive of
    ``output.SnapshotFile``.  Returns (access key, final grid, status)."""
    sv = sim_variables
    now = datetime.now()
    key = now.strftime('%Y%m%d%H%M%S') + str(now.microsecond)
    new = not os.path.exists(path)
    root = {"datetime": now.strftime('%Y%m%d%H%M'), "seed": seed} if new else None
    with open_snapshots(path, "w" if new else "a", attrs=root) as f:
        lap = perf_counter()
        out, st = core_run(grid, sv, precision=precision, save=lambda k, a: f(f"{key}/{k}", a), every=every)
        elapsed = perf_counter() - lap
        f.set_attrs(key, {"config": sv.config, "cells": int(sv.cells), "cfl": float(sv.cfl),
                          "gamma": float(sv.gamma), "dimension": int(sv.dimension), "subgrid": sv.subgrid,
                          "timestep": sv.timestep, "solver": sv.solver, "elapsed": elapsed})
    return key, out, st


def simulate(params, path=None, seed=None, every="step", precision=None):
    """A single run of astrea.py:run() for a parameters mapping: the options
    and the configuration's test conditions (``options.handle_variables``),
    the initial state of constructor.initialise (``ic.initialise``), the
    device time loop, and with ``path`` the snapshot file of ``save_run``.
    Returns (final grid, status, sim_variables)."""
    sv = options.handle_variables(params)
    grid = ic.initialise(sv)
    if path is None:
        out, st = core_run(grid, sv, precision=precision)
    else:
        _, out, st = save_run(path, grid, sv, seed, every, precision)
    return out, st, sv

"""Benchmark workloads (2D shear layer, ragged 1D Riemann ensemble, served
four-quadrant stream) as reusable runners.

Each runner builds its synthetic input on the device from a stated seed,
runs the device update and returns a dict of measurements.  bench.py prints
them; tests use small instances.
"""
import math
import time

import numpy as np
import torch

from . import batch, ic
from .plan import Plan

SEED = 20261202


def rel_l1(a, b):
    """Max over variables of the relative L1 error of a against b (packed
    state planes, or reference-layout arrays)."""
    if isinstance(a, torch.Tensor):
        num = (a - b).abs().sum(dim=1)
        den = b.abs().sum(dim=1)
        return float(torch.max(torch.where(den > 0, num / den.clamp_min(1e-300), num)).item())
    ax = tuple(range(a.ndim - 1))
    num = np.abs(a - b).sum(ax)
    den = np.abs(b).sum(ax)
    return float(np.max(np.where(den > 0, num / np.where(den > 0, den, 1), num)))


# ---------------------------------------------------------------------------
# Ragged 1D Riemann ensemble
# ---------------------------------------------------------------------------
def ensemble_inputs(nprob=512, seed=SEED, nmin=256, nmax=8192, t_scale=1.0, host_grids=False):
    """Problems of the seeded ensemble and their initial data: the Riemann
    states (built on the device when a plan starts, ah_plan_init_ic), or with
    ``host_grids`` the host-pipeline grids."""
    specs = ic.ensemble_1d(nprob, seed, nmin, nmax)
    inits, probs = [], []
    for sp in specs:
        if host_grids:
            inits.append(ic.to_device(ic.riemann_1d(sp["n"], sp["left"], sp["right"]), 1.4, "edge", True))
        else:
            inits.append(dict(left=sp["left"], right=sp["right"]))
        probs.append(dict(n0=sp["n"], boundary="edge", dx=1 / sp["n"], gamma=1.4, cfl=.5,
                          t_end=sp["t_end"] * t_scale))
    return inits, probs


def _start(plan, inits):
    """Initial states: Riemann data built on the device, or grids loaded."""
    if inits and isinstance(inits[0], dict):
        for i, d in enumerate(inits):
            plan.init_riemann(i, d["left"], d["right"])
    else:
        plan.load(inits)


def run_ensemble_batched(grids, probs, precision="fp64", chunk=32, use_graph=True, blocks=None, compat="physical",
                         compact_every=4, tile_s=0):
    """The ensemble in one plan, replayed in graph chunks of `chunk` steps with
    the tile lists compacted every `compact_every` chunks (on the 512-problem
    ensemble, 32 and 4 run 7 % faster than 16 and 1: each compaction waits for
    the device, while a finished problem's tiles only cost an early exit)."""
    plan = Plan(1, "ppm", "hllc", "ssprk(3,3)", probs, varmask=0x13, precision=precision, compat=compat,
                tile_s=tile_s)
    if blocks:
        plan.set_blocks(blocks)
    _start(plan, grids)
    plan.reset()
    torch.cuda.synchronize()
    st, secs, nch = batch.run(plan, chunk=chunk, use_graph=use_graph, compact_every=compact_every)
    zu = sum(p["n0"] * s["steps"] for p, s in zip(probs, st))
    out = plan.state.clone()
    offs = plan.offsets
    plan.close()
    return dict(seconds=secs, zu=zu, zu_per_s=zu / secs, chunks=nch, status=st, state=out, offsets=offs)


def run_ensemble_plain(grids, probs, sample=None, chunk=64, compat="physical"):
    """The plain path: one problem at a time, FP64, eager launches, host
    looks at the device once per chunk.  ``sample``: indices to run (the rate
    is what is compared; a stratified sample keeps the run bounded)."""
    idx = range(len(probs)) if sample is None else sample
    zu, secs, st = 0, 0.0, {}
    for i in idx:
        plan = Plan(1, "ppm", "hllc", "ssprk(3,3)", [probs[i]], varmask=0x13, compat=compat)
        _start(plan, [grids[i]])
        plan.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s = plan.run(chunk)[0]
        torch.cuda.synchronize()
        secs += time.perf_counter() - t0
        zu += probs[i]["n0"] * s["steps"]
        st[i] = s
        plan.close()
    return dict(seconds=secs, zu=zu, zu_per_s=zu / secs, status=st)


def autotune(grids, probs, precision, tiles=(0, 115), candidates=(1024, 2048, 4096, 8192), steps=64,
             compat="physical"):
    """Pick the 1D tile shape (``tile_s``: 0 = 243-cell tiles of 256 threads,
    115 = 115-cell tiles of 128 threads) and the persistent grid cap on the
    problem set itself: time a fixed number of steps of the set for every
    pair (each from the initial state).  Returns ((tile_s, blocks), times)."""
    best, times = None, {}
    for t in tiles:
        b, tb = autotune_blocks(grids, probs, precision, candidates, steps, compat, tile_s=t)
        for k, v in tb.items():
            times[(t, k)] = v
            if best is None or v < times[best]:
                best = (t, k)
    return best, times


def autotune_blocks(grids, probs, precision, candidates=(1024, 2048, 4096, 8192), steps=64, compat="physical",
                    tile_s=0):
    """Pick the persistent grid cap on the problem set itself: time a
    fixed number of steps for each candidate (state restored between)."""
    best, times = None, {}
    for b in candidates:
        plan = Plan(1, "ppm", "hllc", "ssprk(3,3)", probs, varmask=0x13, precision=precision, compat=compat,
                    tile_s=tile_s)
        plan.set_blocks(b)
        _start(plan, grids)
        plan.reset()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            plan.step_graph(16, 1)
            s.synchronize()
            t0 = time.perf_counter()
            plan.step_graph(16, steps // 16)
            s.synchronize()
            times[b] = time.perf_counter() - t0
        plan.close()
        if best is None or times[b] < times[best]:
            best = b
    return best, times


# ---------------------------------------------------------------------------
# 2D shear layer (see bench.py)
# ---------------------------------------------------------------------------
def shear_inputs(n, seed=SEED):
    return ic.to_device(ic.shear_layer(n, seed), 1.4, "wrap", True)

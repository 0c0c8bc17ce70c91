"""Ragged batches of independent problems in one plan.

All problems of a batch share one scheme; their cells are packed back to
back, every kernel walks a tile list that covers all of them, and each
problem keeps its own t, dt, step count and termination on the device
(ah_plan_step).  Runs are chunked: a captured hipGraph of `chunk` steps is
replayed, then the host compacts the tile lists to the problems that are
still active (ah_plan_compact), so finished problems cost nothing.
"""
import time

import torch

from .plan import Plan


def run(plan: Plan, chunk=16, use_graph=True, compact_every=1, max_chunks=10 ** 7):
    """Advance every problem of ``plan`` to its end.  Returns (status list,
    seconds, number of chunks)."""
    caller = torch.cuda.current_stream(plan.device)
    stream = torch.cuda.Stream(device=plan.device) if use_graph else caller
    # the plan's state was loaded on the caller's stream
    stream.wait_stream(caller)
    t0 = time.perf_counter()
    n = 0
    with torch.cuda.stream(stream):
        active = plan.compact()
        while active and n < max_chunks:
            if use_graph:
                plan.step_graph(chunk, 1)
            else:
                plan.step(chunk)
            n += 1
            if n % compact_every == 0:
                active = plan.compact()
        torch.cuda.current_stream(plan.device).synchronize()
    return plan.status(), time.perf_counter() - t0, n

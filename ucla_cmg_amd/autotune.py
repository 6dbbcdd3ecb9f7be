"""Launch-shape autotuner for a live CardiacSolver.

Times each candidate (total CG grid, captured step on/off) over a few real
time steps with HIP events on the solver's stream and keeps the fastest.
The tuning steps advance the simulation like any other steps, so the tuner
belongs in the warm-up of a run."""
from __future__ import annotations

import itertools
from typing import Iterable, Optional

DEFAULT_BLOCKS = (0, 512, 1536, 2048)


def tune(solver, steps_each: int = 10, blocks: Iterable[int] = DEFAULT_BLOCKS,
         graphs: Iterable[bool] = (False, True), step_fn: Optional[callable] = None,
         margin: float = 0.05) -> dict:
    """Returns {"cg_blocks", "graph", "ms_per_step", "table"} and leaves the
    solver on the chosen shape: the first candidate (the library's default
    grid, host-driven) unless another is faster by more than ``margin`` (the
    trial windows are short and the work per step drifts with the activity,
    so smaller differences are noise).  ``step_fn(k)`` advances k steps
    (default ``solver.Solve``); a captured candidate gets extra untimed steps
    first so its host-driven learning steps and capture are not timed."""
    import torch

    step = step_fn or solver.Solve
    stream = torch.cuda.ExternalStream(solver.stream())
    cands = list(itertools.product(blocks, graphs))
    times = {c: [] for c in cands}
    used = {}
    # Two rounds, the second in reverse order: the work per step drifts as
    # the simulation evolves (CG iterations follow the wavefront), and the
    # mean over mirrored positions cancels that drift to first order.
    for order in (cands, cands[::-1]):
        for c in order:
            g, graph = c
            solver.setLaunch(cg_blocks=g, graph=graph)
            step(4 if graph else 1)
            solver.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            step(steps_each)
            e1.record(stream)
            solver.synchronize()
            e1.synchronize()
            times[c].append(e0.elapsed_time(e1) / steps_each)
            used[c] = solver.cgBlocks()
    table = [{"cg_blocks": used[c], "graph": bool(c[1]), "ms_per_step": sum(t) / len(t)}
             for c, t in times.items()]
    best = min(table, key=lambda r: r["ms_per_step"])
    if best["ms_per_step"] > (1.0 - margin) * table[0]["ms_per_step"]:
        best = table[0]
    solver.setLaunch(cg_blocks=best["cg_blocks"], graph=best["graph"])
    return dict(best, table=table)

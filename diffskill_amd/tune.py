"""Launch-shape autotuner: particles per work item and persistent workgroup
counts of the particle (tile) and grid phases, picked by timing the caller's
own forward + adjoint pass under each candidate shape.

The shape changes only scheduling and float summation order, never the
algorithm, so any candidate is a valid choice; the tuner keeps the fastest.
It must run before a pass is captured into a graph (the graph freezes the
launch shapes it recorded).
"""
import time

import torch


def default_candidates(n_cu):
    """(particles_per_item, tile_ctas, grid_ctas) shapes to try."""
    return [(ppi, t * n_cu, 8 * n_cu) for ppi in (256, 512, 1024, 2048) for t in (2, 4, 8)]


def autotune(engine, run_pass, candidates=None, reps=2):
    """Time ``run_pass`` (device work on the current stream) once warm plus
    ``reps`` times per candidate; leave the fastest shape set on ``engine``.
    Returns (best shape dict, {candidate: ms per pass})."""
    if candidates is None:
        candidates = default_candidates(torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count)
    timings = {}
    for cand in candidates:
        engine.set_launch(*cand)
        run_pass()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            run_pass()
        torch.cuda.synchronize()
        timings[tuple(cand)] = (time.perf_counter() - t0) * 1000 / reps
    best = min(timings, key=timings.get)
    engine.set_launch(*best)
    return engine.get_launch(), timings

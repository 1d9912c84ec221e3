"""Launch-shape autotuner for a resident GpuSolver.

It picks, one key at a time (coordinate descent, in the order of DEFAULT_SPACE), the osph_tune value with
the shortest measured PC step for the solver's own particles. The step times come from wall-clock
timing of `steps` steps per candidate after one untimed step; the starting settings are run once untimed
before their timed run, since the first steps on a state size the buffers. The solver's state is
downloaded first and uploaded again after every candidate, so the tuning leaves the resident state as
it found it (the derivative planes included).

Every candidate gives the same neighbour sets. The lane split of the pair kernels changes the order of
the summation, so the derivatives may differ in the last bits (within the parity tolerances). No
reference counterpart: the reference has no device kernels.
"""
from __future__ import annotations

import time
from typing import Dict, Iterable, List, Sequence, Tuple

# key -> candidate values, in the order the descent visits them (the first value of each list is the
# library's default)
DEFAULT_SPACE: List[Tuple[str, Sequence[int]]] = [
    ("pairs_lanes", (1, 2, 4)),
    ("cell_divisor", (3, 2, 4)),
    ("lists_block", (256, 128)),
    ("pairs_block", (256, 128)),
    ("fuse_finalize", (-1, 0, 1)),
]


def _time_steps(solver, steps: int) -> float:
    solver.step()
    solver.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.step()
    solver.sync()
    return (time.perf_counter() - t0) / steps


def autotune(solver, steps: int = 5, space: Iterable[Tuple[str, Sequence[int]]] | None = None,
             log=None) -> Dict[str, object]:
    """Tunes `solver` in place; returns {"settings": {key: value}, "ms_per_step": best, "trials": [...]}"""
    space = list(DEFAULT_SPACE if space is None else space)
    saved, saved_dt = solver.download(), solver.dt
    best: Dict[str, int] = {key: values[0] for key, values in space}
    solver.tune(**best)
    _time_steps(solver, steps)  # the first steps on a state size the buffers: kept out of the comparison
    solver.upload(saved)
    solver.dt = saved_dt
    best_t = _time_steps(solver, steps)
    solver.upload(saved)
    solver.dt = saved_dt
    trials = [(dict(best), best_t)]
    if log:
        log(f"autotune {best}: {1e3 * best_t:.3f} ms/step")
    for key, values in space:
        for value in values[1:]:
            trial = dict(best)
            trial[key] = value
            solver.tune(**trial)
            t = _time_steps(solver, steps)
            solver.upload(saved)
            solver.dt = saved_dt
            trials.append((trial, t))
            if log:
                log(f"autotune {trial}: {1e3 * t:.3f} ms/step")
            if t < best_t:
                best, best_t = trial, t
        solver.tune(**best)
    return {"settings": best, "ms_per_step": 1e3 * best_t, "trials": trials}

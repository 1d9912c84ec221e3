"""Launch-shape autotuner for a resident GpuSolver: the Python face of osph_autotune (the tuner itself is
C++ in the library, so a C++ host gets the same search).

It picks, one key at a time (coordinate descent, in the order of DEFAULT_SPACE), the osph_tune value with
the shortest measured PC step for the solver's own particles. Each candidate is timed over `steps` steps
after one untimed step; the starting settings are run once before their timed run, since the first steps
on a state size the buffers. The resident state is copied on the device before the first candidate and
restored after every one, so the tuning leaves the run where it was (the derivative planes included).

Every candidate gives the same neighbour sets. The lane split of the pair kernels changes the order of
the summation, so the derivatives may differ in the last bits (within the parity tolerances). No
reference counterpart: the reference has no device kernels.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Iterable, List, Sequence, Tuple

# key -> candidate values, in the order the descent visits them (the first value of each list is the
# library's default); the library's own default space is the same
DEFAULT_SPACE: List[Tuple[str, Sequence[int]]] = [
    ("pairs_lanes", (1, 2, 4)),
    ("cell_divisor", (3, 2, 4)),
    ("lists_block", (256, 128)),
    ("pairs_block", (256, 128)),
    ("fuse_finalize", (1, 0, -1)),
]


def autotune(solver, steps: int = 5, space: Iterable[Tuple[str, Sequence[int]]] | None = None,
             log=None) -> Dict[str, object]:
    """Tunes `solver` in place (osph_autotune); returns {"settings": {key: value}, "ms_per_step": best,
    "trials": [(settings, seconds per step), ...]} in the order the candidates ran."""
    space = list(DEFAULT_SPACE if space is None else space)
    n = len(space)
    keys = (C.c_char_p * n)(*[k.encode() for k, _ in space])
    counts = (C.c_int * n)(*[len(v) for _, v in space])
    flat = [int(x) for _, v in space for x in v]
    values = (C.c_longlong * len(flat))(*flat)
    chosen = (C.c_longlong * n)()
    cap = 1 + sum(len(v) - 1 for _, v in space)
    trial_ms = (C.c_double * cap)()
    best_ms = C.c_double()
    count = C.c_int()
    solver._check(solver.L.osph_autotune(solver.ctx, float(solver.dt), int(steps), keys, counts, values, n, chosen,
                                         C.byref(best_ms), trial_ms, cap, C.byref(count)), "osph_autotune")
    settings = {k: int(chosen[i]) for i, (k, _) in enumerate(space)}
    # the candidates in the order they ran: the start, then each key's other values from the best so far
    trials, cur = [], {k: v[0] for k, v in space}
    order = [dict(cur)]
    best_t, j = trial_ms[0], 1
    for key, vals in space:
        for v in vals[1:]:
            trial = dict(cur)
            trial[key] = v
            order.append(trial)
            if trial_ms[j] < best_t:
                best_t = trial_ms[j]
                cur = trial
            j += 1
    for k in range(min(count.value, cap)):
        trials.append((order[k], trial_ms[k] * 1e-3))
        if log:
            log(f"autotune {order[k]}: {trial_ms[k]:.3f} ms/step")
    return {"settings": settings, "ms_per_step": best_ms.value, "trials": trials}

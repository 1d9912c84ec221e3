"""The launch-shape autotuner (osph_autotune): the chosen settings come from the search space, the resident
state is the one it found, and the tuned solver steps exactly like an untuned twin given the same settings."""
import numpy as np
import pytest

from opensph_amd import autotune, gpu

pytestmark = pytest.mark.gpu

SPACE = [("pairs_lanes", (1, 2, 4)), ("cell_divisor", (3, 2)), ("fuse_finalize", (-1, 1))]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotune_keeps_state_and_picks_from_space(precision):
    from cases import exercised_fluid

    st, cfg = exercised_fluid()
    print(f"[case] autotune {precision} n={st.n}", flush=True)
    g = gpu.GpuSolver(st, cfg, precision=precision)
    g.integrate()
    twin = gpu.GpuSolver(st, cfg, precision=precision)  # the same history, never tuned
    twin.integrate()
    before, dt0 = g.download(st), g.dt
    res = autotune.autotune(g, steps=2, space=SPACE)
    after = g.download(st)
    for k in ("r", "v", "rho", "u", "dv", "drho", "du"):
        assert np.array_equal(after[k], before[k]), k
    assert g.dt == dt0
    space = dict(SPACE)
    assert set(res["settings"]) == set(space)
    for k, v in res["settings"].items():
        assert v in space[k], k
    assert len(res["trials"]) == 1 + sum(len(v) - 1 for _, v in SPACE)
    assert res["ms_per_step"] > 0.0

    # the tuning left the run where it was (state, slot order, search reference): the tuned solver and an
    # untuned twin with the same history, given the chosen settings, take identical steps
    twin.tune(**res["settings"])
    for _ in range(3):
        assert g.step() == twin.step()
    a, b = g.download(st), twin.download(st)
    for k in ("r", "v", "rho", "u"):
        assert np.array_equal(a[k], b[k]), k
    g.close()
    twin.close()

"""Checkpoint and resume through the reference's .ssf format: a run saved after some steps and loaded into
a new solver continues exactly as the uninterrupted run (the resident state is never reordered here, so
both runs sum every particle's neighbours in the same order)."""
import numpy as np
import pytest

from opensph_amd import gpu, ssf

pytestmark = pytest.mark.gpu

STATE = ("r", "v", "dv", "rho", "drho", "u", "du", "s", "ds", "damage", "ddamage", "divv")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_resume_from_ssf_continues_the_run(precision, tmp_path):
    from cases import exercised_impact

    st, cfg = exercised_impact()
    print(f"[case] ssf resume {precision} n={st.n}", flush=True)
    a = gpu.GpuSolver(st, cfg, precision=precision)
    a.tune(reorder_period=0)
    for _ in range(4):
        a.step()
    path = tmp_path / "checkpoint.ssf"
    a.save(path, run_time=1.0)
    assert ssf.read(path).time_step == a.dt
    b = gpu.GpuSolver.load(path, cfg, precision=precision)
    b.tune(reorder_period=0)
    la = [a.step() for _ in range(3)]
    lb = [b.step() for _ in range(3)]
    assert la == lb
    x, y = a.download(st), b.download(st)
    for k in STATE:
        assert np.array_equal(x[k], y[k]), k
    a.close()
    b.close()

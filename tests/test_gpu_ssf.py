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


SAVED = ("r", "v", "dv", "rho", "drho", "u", "du", "s", "ds", "damage", "ddamage")


def _compare_with_oracle(got, ref, tol, state_tol=None):
    from conftest import rel_l2

    for k in SAVED:
        x, y = got[k], ref[k]
        if k in ("v", "dv"):
            x, y = x[:, :3], y[:, :3]
        bar = state_tol if state_tol is not None and k in ("r", "v", "rho") else tol
        assert rel_l2(x, y) <= bar, (k, rel_l2(x, y))
    for k in ("flag", "mat"):
        assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("precision,tol,state_tol", [("fp64", 1e-10, None), ("mixed", 1e-5, 1e-8)])
def test_ssf_after_k_steps_holds_the_oracle_state(precision, tol, state_tol, oracle, tmp_path):
    """The GPU run saved after K steps (BinaryOutput's layout, opensph_amd.ssf) holds the state of the oracle
    after K PredictorCorrector steps; the file's time step is the oracle's next time step.  Mixed precision:
    the rates at the acceleration bar, positions, velocities and densities (FP64 state integrating FP32
    rates over K steps) at 1e-8."""
    from cases import exercised_impact

    st, cfg = exercised_impact()
    K = 5
    print(f"[case] ssf after {K} steps {precision} n={st.n}", flush=True)
    a = gpu.GpuSolver(st, cfg, precision=precision)
    for _ in range(K):
        a.step()
    path = tmp_path / "k.ssf"
    a.save(path, run_time=1.0)
    a.close()
    f = ssf.read(path)
    ref, log = oracle.run_pc(st, cfg, K)
    assert abs(f.time_step - log[-1][1]) <= tol * log[-1][1]
    assert f.particle_count == st.n
    _compare_with_oracle(ssf.to_state(f), ref, tol, state_tol)


def test_cpp_checkpoint_through_reference_binary_output(oracle, tmp_path):
    """A C++ host checkpoints through the adapter: GpuSymmetricSolver::sync, then the reference's own
    BinaryOutput::dump (core/io/Output.cpp:481-595) of its Storage.  The file holds the oracle's state after
    the same steps, and BinaryInput's layout reads back through opensph_amd.ssf."""
    from cases import exercised_impact
    from oracle.bridge import AdapterRun, RefRun

    if not (AdapterRun.available() and RefRun.available()):
        pytest.skip("the adapter library (oracle/_ref/libosph_adapter.so) was not built")
    st, cfg = exercised_impact()
    K = 4
    print(f"[case] adapter checkpoint after {K} steps n={st.n}", flush=True)
    run = AdapterRun(st, cfg)
    for _ in range(K):
        _, nxt, _ = run.step()
    path = tmp_path / "cpp.ssf"
    run.save(path, run_time=2.0)
    run.close()
    f = ssf.read(path)
    assert f.run_time == 2.0 and f.time_step == nxt
    ref, log = oracle.run_pc(st, cfg, K)
    assert abs(nxt - log[-1][1]) <= 1e-10 * log[-1][1]
    _compare_with_oracle(ssf.to_state(f), ref, 1e-10)

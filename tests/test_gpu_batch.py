"""Ensembles: osph_step_pc_batch steps several independent resident runs phase by phase (each with its own
time step).  Every member must end bit-identical to the same run stepped alone with osph_step_pc."""
import numpy as np
import pytest

from opensph_amd import gpu, initial

pytestmark = pytest.mark.gpu

FIELDS = ("r", "v", "dv", "rho", "drho", "u", "du", "s", "ds", "damage", "p", "cs", "divv", "gradv", "ncnt")


def members():
    out = []
    for k, (nt, npj, ang, speed) in enumerate([(6000, 600, 30.0, 3e3), (3000, 300, 75.0, 6e3), (9000, 500, 10.0, 7e3)]):
        st, cfg = initial.impact(n_target=nt, n_projectile=npj, r_projectile=0.2 * 50e3 * (npj / nt) ** (1 / 3),
                                 angle_deg=ang, speed=speed, offset_h=-0.3, seed=100 + k)
        out.append((st, cfg, "fp64"))
    st, cfg = initial.impact(n_target=4000, n_projectile=400, offset_h=-0.3, seed=7)
    out.append((st, cfg, "mixed"))
    st, cfg = initial.fluid_sphere(n_requested=2000, radius=1000.0)
    out.append((st, cfg, "fp64"))
    return out


def test_batch_steps_equal_single_steps():
    ms = members()
    print(f"[case] batch of {len(ms)}: " + ", ".join(str(st.n) for st, _, _ in ms), flush=True)
    batch = [gpu.GpuSolver(st, cfg, precision=p) for st, cfg, p in ms]
    alone = [gpu.GpuSolver(st, cfg, precision=p) for st, cfg, p in ms]
    for step in range(6):
        nxt, crit = gpu.step_batch(batch)
        for k, s in enumerate(alone):
            _, n1, c1 = s.step()
            assert n1 == nxt[k] and c1 == crit[k], (step, k)
    for k, ((st, _, _), a, b) in enumerate(zip(ms, batch, alone)):
        da, db = a.download(st), b.download(st)
        for f in FIELDS:
            assert np.array_equal(da[f], db[f]), (k, f)


def test_batch_with_an_empty_member_and_errors():
    st, cfg = initial.fluid_sphere(n_requested=500, radius=100.0)
    empty = st.subset(np.zeros(0, dtype=np.int64))
    print("[case] batch with empty member", flush=True)
    a, b = gpu.GpuSolver(st, cfg), gpu.GpuSolver(empty, cfg)
    nxt, _ = gpu.step_batch([a, b])
    assert nxt[0] > 0 and nxt[1] == cfg.max_dt
    with pytest.raises(gpu.GpuError):
        gpu.step_batch([a, a])


def test_batch_with_self_gravity_equals_single_steps():
    """Members with an attached self-gravity (Barnes-Hut and the direct sum, FP64 and mixed) step in a
    batch exactly as alone: each member's gravity is evaluated on its own predicted state."""
    import dataclasses

    ms = members()[:2] + members()[3:4]
    grav = [2, 1, 2]
    print(f"[case] batch with gravity: " + ", ".join(str(st.n) for st, _, _ in ms), flush=True)

    def make():
        return [gpu.GpuSolver(st, dataclasses.replace(cfg, gravity=gv), precision=p)
                for (st, cfg, p), gv in zip(ms, grav)]

    batch, alone = make(), make()
    for step in range(4):
        nxt, crit = gpu.step_batch(batch)
        for k, s in enumerate(alone):
            _, n1, c1 = s.step()
            assert n1 == nxt[k] and c1 == crit[k], (step, k)
    for k, ((st, _, _), a, b) in enumerate(zip(ms, batch, alone)):
        da, db = a.download(st), b.download(st)
        for f in FIELDS:
            assert np.array_equal(da[f], db[f]), (k, f)
    for s in batch + alone:
        s.close()


def homogeneous_members(precision):
    """Impacts of several sizes and geometries, one fluid sphere whose smoothing lengths give more than the
    default 96 neighbours (its lists overflow on the first step and are grown), and an empty state."""
    out = []
    for k, (nt, npj, ang, speed) in enumerate([(6000, 600, 30.0, 3e3), (3000, 300, 75.0, 6e3), (9000, 500, 10.0, 7e3),
                                               (1500, 200, 50.0, 4e3)]):
        st, cfg = initial.impact(n_target=nt, n_projectile=npj, r_projectile=0.2 * 50e3 * (npj / nt) ** (1 / 3),
                                 angle_deg=ang, speed=speed, offset_h=-0.3, seed=200 + k)
        out.append((st, cfg))
    st, cfg = initial.fluid_sphere(n_requested=3000, radius=1000.0, sigma_v=10.0)
    st["r"][:, 3] *= 1.7
    out.append((st, cfg))
    out.insert(2, (st.subset(np.zeros(0, dtype=np.int64)), cfg))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("graph", [1, 0])
def test_one_launch_per_stage_equals_single_steps(precision, graph):
    """Members of one precision step with one launch per stage for the whole batch (the stages after the
    sort replayed from a captured HIP graph, or launched eagerly): over 12 steps, which cross the reorder
    into cell order and a list overflow, every member is bit-identical to stepping it alone."""
    ms = homogeneous_members(precision)
    print(f"[case] one launch per stage {precision} graph={graph}: " + ", ".join(str(st.n) for st, _ in ms),
          flush=True)
    batch = [gpu.GpuSolver(st, cfg, precision=precision) for st, cfg in ms]
    alone = [gpu.GpuSolver(st, cfg, precision=precision) for st, cfg in ms]
    batch[0].tune(batch_graph=graph)
    for step in range(12):
        nxt, crit = gpu.step_batch(batch)
        for k, s in enumerate(alone):
            _, n1, c1 = s.step()
            assert n1 == nxt[k] and c1 == crit[k], (step, k)
    for k, ((st, _), a, b) in enumerate(zip(ms, batch, alone)):
        da, db = a.download(st), b.download(st)
        for f in FIELDS:
            assert np.array_equal(da[f], db[f]), (k, f)
    # single steps after batched ones continue the same run
    for s, t in zip(batch, alone):
        s.step()
        t.step()
    for k, ((st, _), a, b) in enumerate(zip(ms, batch, alone)):
        assert np.array_equal(a.download(st)["r"], b.download(st)["r"]), k
    for s in batch + alone:
        s.close()

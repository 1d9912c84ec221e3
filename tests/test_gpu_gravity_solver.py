"""SPH with self-gravity in one step (osph_attach_gravity; GravitySolver<SymmetricSolver<3>>,
core/sph/solvers/GravitySolver.cpp:64-91) against the oracle's composition of the same: the direct sum
added to the accelerations, then the pair sums.  Bars as for the plain path (1e-10 relative per step);
Barnes-Hut with every node opened is the direct sum; the mixed path within its acceleration bar."""
import dataclasses

import numpy as np
import pytest

from conftest import rel_l2
from opensph_amd import gpu

pytestmark = pytest.mark.gpu

TOL = 1e-10
DERIV = ("dv", "du", "divv", "drho", "p", "cs", "ds", "gradv", "ddamage")
STATE = ("r", "v", "dv", "rho", "drho", "u", "du", "s", "ds", "damage", "ddamage")


def _close(a, b, fields, tol=TOL):
    for k in fields:
        x, y = a[k], b[k]
        if k == "dv":
            x, y = x[:, :3], y[:, :3]
        assert rel_l2(x, y) < tol, (k, rel_l2(x, y))


def _cases():
    from cases import exercised_fluid, exercised_impact

    return {"fluid_exercised": exercised_fluid, "impact_exercised": exercised_impact}


@pytest.mark.parametrize("name", ["fluid_exercised", "impact_exercised"])
@pytest.mark.parametrize("solver", [1, 2])
def test_integrate_with_gravity_matches_oracle(oracle, name, solver):
    st, cfg = _cases()[name]()
    cfg = dataclasses.replace(cfg, gravity=solver, gravity_theta=1e-12)  # Barnes-Hut opens every node
    print(f"[case] gravity solver {name} solver={solver} n={st.n}", flush=True)
    g = gpu.GpuSolver(st, cfg)
    g.integrate()
    a = g.download(st)
    b = oracle.integrate(st, dataclasses.replace(cfg, gravity=1))
    _close(a, b, DERIV)
    plain = oracle.integrate(st, dataclasses.replace(cfg, gravity=0))
    grav = b["dv"][:, :3] - plain["dv"][:, :3]
    assert np.linalg.norm(grav) > 10 * TOL * np.linalg.norm(b["dv"][:, :3])  # a missing term would fail
    g.close()


def test_steps_with_gravity_match_oracle(oracle):
    st, cfg = _cases()["impact_exercised"]()
    cfg = dataclasses.replace(cfg, gravity=1, criteria=cfg.criteria | 4)
    print(f"[case] gravity solver steps n={st.n}", flush=True)
    g = gpu.GpuSolver(st, cfg)
    glog = [g.step() for _ in range(5)]
    so, olog = oracle.run_pc(st, cfg, 5)
    for (dt, nxt, crit), (odt, onxt, ocrit) in zip(glog, olog):
        assert nxt == pytest.approx(onxt, rel=TOL)
        assert crit == ocrit
    _close(g.download(st), so, STATE)
    g.close()


def test_dominant_gravity_mixed_against_fp64():
    """Barnes-Hut (theta 0.5, octupole) with G raised so that gravity dominates the accelerations: the
    mixed path (FP32 pair sums and FP32 list evaluation) within 1e-5 of the FP64 path."""
    st, cfg = _cases()["fluid_exercised"]()
    cfg = dataclasses.replace(cfg, gravity=2, gravity_G=1.0)
    print(f"[case] gravity solver mixed n={st.n}", flush=True)
    out = {}
    for prec in ("fp64", "mixed"):
        g = gpu.GpuSolver(st, cfg, precision=prec)
        g.integrate()
        out[prec] = g.download(st)
        g.close()
    plain = gpu.GpuSolver(st, dataclasses.replace(cfg, gravity=0))
    plain.integrate()
    p = plain.download(st)
    plain.close()
    grav = out["fp64"]["dv"][:, :3] - p["dv"][:, :3]
    assert np.linalg.norm(grav) > 0.5 * np.linalg.norm(out["fp64"]["dv"][:, :3])
    assert rel_l2(out["mixed"]["dv"][:, :3], out["fp64"]["dv"][:, :3]) < 1e-5
    assert np.array_equal(out["mixed"]["ncnt"], out["fp64"]["ncnt"])


def test_detached_gravity_gives_the_plain_solver():
    st, cfg = _cases()["fluid_exercised"]()
    print(f"[case] gravity solver detach n={st.n}", flush=True)
    a = gpu.GpuSolver(st, dataclasses.replace(cfg, gravity=2))
    grav = a.gravity
    a.attach_gravity(None)
    la = [a.step() for _ in range(3)]
    b = gpu.GpuSolver(st, cfg)
    lb = [b.step() for _ in range(3)]
    assert la == lb
    x, y = a.download(st), b.download(st)
    for k in STATE:
        assert np.array_equal(x[k], y[k]), k
    a.close()
    b.close()
    grav.close()


def test_gravity_on_empty_and_single_particle_states():
    from opensph_amd import initial

    st, cfg = initial.fluid_sphere(n_requested=200, radius=100.0)
    cfg = dataclasses.replace(cfg, gravity=2)
    for idx in ([], [0]):
        sub = st.subset(np.asarray(idx, dtype=np.int64))
        print(f"[case] gravity solver n={sub.n}", flush=True)
        g = gpu.GpuSolver(sub, cfg)
        g.integrate()
        dt, nxt, crit = g.step()
        assert np.isfinite(nxt)
        if sub.n:
            assert np.all(g.download(sub)["dv"][:, :3] == 0.0)  # no other particle, no force
        g.close()


def test_gravity_context_serves_one_solver():
    """A gravity context holds one SPH context's particles and accelerations between its evaluation and
    that context's pair sums, so attaching it to a second context fails until the first lets it go."""
    from opensph_amd.gravity import GpuGravity

    st, cfg = _cases()["fluid_exercised"]()
    print(f"[case] gravity ownership n={st.n}", flush=True)
    a = gpu.GpuSolver(st, cfg)
    b = gpu.GpuSolver(st, cfg)
    g = GpuGravity(st["r"], st["m"], solver="brute_force")
    a.attach_gravity(g)
    with pytest.raises(gpu.GpuError, match="another SPH context"):
        b.attach_gravity(g)
    a.attach_gravity(None)
    b.attach_gravity(g)  # free again after the detach
    b.integrate()
    b.attach_gravity(None)
    a.attach_gravity(g)
    a.gravity = None  # the Python side must not close g with its owner
    a.close()  # destroying the owner releases it too
    c = gpu.GpuSolver(st, cfg)
    c.attach_gravity(g)
    c.attach_gravity(None)
    for s in (b, c):
        s.gravity = None
        s.close()
    g.close()


def test_gravity_context_moved_by_leapfrog_then_attached():
    """A gravity context whose state LeapFrog steps moved into Morton order, then attached to an SPH solver:
    it loads the solver's particles in the solver's order, and the SPH + Barnes-Hut steps give the same
    bits as with a fresh gravity context."""
    from opensph_amd.gravity import GpuGravity

    st, cfg = _cases()["impact_exercised"]()
    cfg = dataclasses.replace(cfg, gravity=2, gravity_G=1.0)
    print(f"[case] gravity moved then attached n={st.n}", flush=True)
    fresh = gpu.GpuSolver(st, cfg)
    plain = gpu.GpuSolver(st, dataclasses.replace(cfg, gravity=0))
    rng = np.random.default_rng(5)
    n = st.n + 2000  # more than the solver's particles: its planes are reused, not reallocated
    r = np.c_[rng.uniform(-1e4, 1e4, (n, 3)), np.full(n, 50.0)]
    g = GpuGravity(r, np.full(n, 1e10), theta=cfg.gravity_theta, order=cfg.gravity_order,
                   leaf_size=cfg.gravity_leaf, G=cfg.gravity_G, precision="fp64")
    g.leapfrog(1.0, 3)  # the first build moves its state into Morton order
    plain.attach_gravity(g)
    for _ in range(3):
        assert plain.step() == fresh.step()
    a, b = plain.download(st), fresh.download(st)
    for k in STATE:
        assert np.array_equal(a[k], b[k]), k
    plain.close()
    fresh.close()

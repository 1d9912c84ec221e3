"""Initial conditions built on the device (osph_setup_spheres) against the host construction of
opensph_amd.initial, which follows HexagonalPacking::generate (core/sph/initial/Distribution.cpp:126-201),
InitialConditions::setQuantities (core/sph/initial/Initial.cpp:293-333) and CollisionGeometrySetupJob
(core/run/jobs/ParticleJobs.cpp:532-577); the host construction's particle counts are pinned against the
reference in test_initial.py.  Every quantity the set-up writes must be bit-identical."""
import numpy as np
import pytest

from opensph_amd import gpu, initial

pytestmark = pytest.mark.gpu

FIELDS = ("r", "v", "m", "rho", "u", "flag", "mat", "reduce")


def assert_same_bodies(solver, st):
    out = solver.download()
    assert out.n == st.n
    for k in FIELDS:
        assert np.array_equal(out[k], st[k]), k
    for a, b in zip(solver.materials, st.materials):
        assert a["rho_lo"] == b["rho_lo"]


@pytest.mark.parametrize("kw", [
    dict(n_target=20_000, n_projectile=1_500, r_target=50e3, r_projectile=5e3, speed=5e3, angle_deg=45.0),
    dict(n_target=3_000, n_projectile=400, r_target=1e3, r_projectile=300.0, speed=3e3, angle_deg=0.0,
         offset_h=-0.5),
    dict(n_target=5_000, n_projectile=100, r_target=50e3, r_projectile=2.6e3, speed=7e3, angle_deg=89.0),
])
def test_device_impact_equals_host_impact(kw):
    print(f"[case] impact {kw}", flush=True)
    st, _ = initial.impact(flaws=False, **kw)
    bodies, mats, cfg, col = initial.impact_setup(**kw)
    solver, info = gpu.GpuSolver.from_bodies(bodies, mats, cfg, collision=col)
    assert [(f, c) for f, c, _, _ in info] == [(f, c) for f, c in initial.bodies(st)]
    for (f, c, vol, h), material in zip(info, st.materials):
        assert vol == float(np.cumsum(st["m"][f:f + c] / st["rho"][f:f + c])[-1])
        assert h == st["r"][f, 3]
    assert_same_bodies(solver, st)
    solver.close()


def test_device_single_sphere_equals_host():
    mat = initial.basalt()
    r, m = initial.sphere_body(7_000, 800.0, mat, center=(10.0, -5.0, 2.5), eta=1.5)
    solver, info = gpu.GpuSolver.from_bodies(
        [dict(particle_count=7_000, radius=800.0, center=(10.0, -5.0, 2.5), eta=1.5, flag=3)], [mat],
        initial.RunConfig())
    out = solver.download()
    assert info[0][:2] == (0, r.shape[0])
    assert np.array_equal(out["r"], r)
    assert np.array_equal(out["m"], m)
    assert np.all(out["flag"] == 3)
    assert solver.materials[0]["rho_lo"] == initial.density_lower_bound(r, m, mat)
    solver.close()


def test_device_ensemble_members_equal_host_members():
    specs = initial.ensemble_specs(6, seed=3, n_range=(2e3, 3e4))
    for spec in specs:
        print(f"[case] ensemble member {spec['n_target']} + {spec['n_projectile']}", flush=True)
        st, _ = initial.impact(flaws=False, **spec)
        bodies, mats, cfg, col = initial.impact_setup(**spec)
        solver, _ = gpu.GpuSolver.from_bodies(bodies, mats, cfg, precision="mixed", collision=col)
        assert_same_bodies(solver, st)
        solver.close()


def test_device_flaws_after_device_setup_equal_upload_path():
    kw = dict(n_target=12_000, n_projectile=900, r_target=50e3, r_projectile=5e3, speed=5e3, angle_deg=30.0)
    st, cfg = initial.impact(flaws=False, **kw)
    up = gpu.GpuSolver(st, cfg)
    initial.device_flaws(up, st, seed=40)
    want = up.download()
    dev = initial.device_impact(flaw_seed=40, **kw)
    got = dev.download()
    for k in ("eps_min", "m_zero", "growth", "n_flaws") + FIELDS:
        assert np.array_equal(got[k], want[k]), k
    # and the two step alike
    for _ in range(3):
        a = up.step()
        b = dev.step()
        assert a == b
    x, y = up.download(), dev.download()
    for k in ("r", "v", "dv", "rho", "u", "s", "damage"):
        assert np.array_equal(x[k], y[k]), k
    up.close()
    dev.close()


@pytest.mark.slow
def test_device_full_impact_equals_host():
    """The 1,062,568-particle impact (950k + 50k requested), bit for bit."""
    st, _ = initial.impact(flaws=False)
    bodies, mats, cfg, col = initial.impact_setup()
    solver, info = gpu.GpuSolver.from_bodies(bodies, mats, cfg, precision="mixed", collision=col)
    assert st.n == 1_062_568
    assert_same_bodies(solver, st)
    solver.close()


def test_setup_argument_errors():
    mat = initial.basalt()
    solver = gpu.GpuSolver(initial.ParticleState(0, [mat]), initial.RunConfig())
    with pytest.raises(gpu.GpuError):
        solver.setup_spheres([dict(particle_count=100, radius=-1.0)])
    with pytest.raises(gpu.GpuError):
        solver.setup_spheres([dict(particle_count=100, radius=10.0, material=3)])
    with pytest.raises(gpu.GpuError):
        solver.setup_spheres([dict(particle_count=0, radius=10.0)])
    assert solver.setup_spheres([]) == []
    solver.close()

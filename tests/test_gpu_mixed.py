"""The mixed-precision path (FP32 pair sums on fixed-point positions and velocities, accumulated in FP64;
FP64 state and integration): derivative error against the FP64 path on the same state, exact neighbour counts, and
energy conservation over 100 predictor-corrector steps."""
import numpy as np
import pytest

from conftest import rel_l2
from opensph_amd import gpu, initial

pytestmark = pytest.mark.gpu

ACCEL_TOL = 1e-5


def total_energy(st):
    """TotalEnergy without gravity (core/physics/Integrals.cpp): sum m (v^2/2 + u)."""
    v2 = np.sum(st["v"][:, :3] ** 2, axis=1)
    return float(np.sum(st["m"] * (0.5 * v2 + st["u"])))


@pytest.mark.parametrize("name,lanes", [("impact_exercised", 1), ("fluid_exercised", 1), ("impact_exercised", 2),
                                        ("impact_exercised", 4)])
def test_mixed_derivatives_match_fp64(name, lanes):
    from cases import exercised_fluid, exercised_impact

    st, cfg = {"impact_exercised": exercised_impact, "fluid_exercised": exercised_fluid}[name]()
    print(f"[case] mixed {name} lanes={lanes} n={st.n}", flush=True)
    a = gpu.GpuSolver(st, cfg, precision="fp64")
    a.integrate()
    da = a.download(st)
    b = gpu.GpuSolver(st, cfg, precision="mixed")
    b.tune(pairs_lanes=lanes)
    b.integrate()
    db = b.download(st)
    assert np.array_equal(da["ncnt"], db["ncnt"])
    assert rel_l2(db["dv"][:, :3], da["dv"][:, :3]) < ACCEL_TOL
    for k in ("divv", "drho", "du", "gradv", "ds"):
        assert rel_l2(db[k], da[k]) < 1e-4, k


@pytest.mark.parametrize("flags", ["codes", "boundary", "wide", "shifted"])
def test_mixed_groups_and_frames_match_fp64(flags):
    """SUM_ONLY_UNDAMAGED groups in the mixed records (4-bit codes, and the group array beyond them) and the
    fixed-point frames far from the origin at a high bulk speed: the mixed derivatives match the FP64
    path, whose group test compares the flags themselves.  A group compared wrongly adds or drops whole
    velocity-gradient terms, far beyond the tolerance."""
    from cases import exercised_impact

    st, cfg = exercised_impact()
    body = st["flag"].copy()
    rng = np.random.default_rng(11)
    if flags == "codes":
        st["flag"][:] = np.where(body == 0, 3, 9)
    elif flags == "boundary":  # 13 is the last coded group, 14 the first one compared through the array
        st["flag"][:] = np.where(body == 0, 13, 14)
    elif flags == "wide":  # groups beyond the codes next to coded ones, split inside the target
        st["flag"][:] = np.where(body == 0, np.where(rng.random(st.n) < 0.5, 500, 2), 1 << 20)
    else:
        st["r"][:, :3] += np.array([3.0e7, -2.0e7, 1.0e7])
        st["v"][:, :3] += np.array([2.0e4, -3.0e4, 1.0e4])
    assert cfg.sum_only_undamaged and cfg.solid_stress
    print(f"[case] mixed groups/frames {flags} n={st.n}", flush=True)
    a = gpu.GpuSolver(st, cfg, precision="fp64")
    a.integrate()
    da = a.download(st)
    b = gpu.GpuSolver(st, cfg, precision="mixed")
    b.integrate()
    db = b.download(st)
    assert np.array_equal(da["ncnt"], db["ncnt"])
    assert rel_l2(db["dv"][:, :3], da["dv"][:, :3]) < ACCEL_TOL
    for k in ("divv", "drho", "du", "gradv", "ds"):
        assert rel_l2(db[k], da[k]) < 1e-4, k


def test_mixed_impact_energy_over_100_steps():
    """Small two-body impact in contact: 100 steps in both precisions; the total energy of the mixed run
    stays within 1e-6 (relative) of the FP64 run, and both conserve it to the integrator's accuracy."""
    st, cfg = initial.impact(n_target=40000, n_projectile=4000, r_projectile=10e3, offset_h=-0.5, seed=2)
    print(f"[case] energy n={st.n}", flush=True)
    e0 = total_energy(st)
    out = {}
    for prec in ("fp64", "mixed"):
        g = gpu.GpuSolver(st, cfg, precision=prec)
        for _ in range(100):
            g.step()
        out[prec] = g.download(st)
    e64, emix = total_energy(out["fp64"]), total_energy(out["mixed"])
    assert abs(emix - e64) / e0 < 1e-6, (e0, e64, emix)
    assert abs(e64 - e0) / e0 < 1e-3
    assert rel_l2(out["mixed"]["r"][:, :3], out["fp64"]["r"][:, :3]) < 1e-5


@pytest.mark.parametrize("name", ["impact_exercised", "mixed_h"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("screen", [0, 1])
def test_screened_lists_keep_pairs_exact(oracle, name, precision, screen):
    """The FP32-screened neighbour test (with FP64 re-tests inside its error band) gives the exact pair
    set in both precisions."""
    from cases import exercised_impact

    if name == "mixed_h":
        rng = np.random.default_rng(5)
        st, cfg = initial.fluid_sphere(n_requested=3000, radius=1000.0)
        st["r"][:, 3] *= np.exp2(rng.uniform(-2.0, 2.0, size=st.n))
    else:
        st, cfg = exercised_impact()
    print(f"[case] screened lists {name} {precision} screen={screen} n={st.n}", flush=True)
    g = gpu.GpuSolver(st, cfg, precision=precision)
    g.tune(lists_screen=screen)
    pi, pj = g.pairs()
    oi, oj = oracle.pairs(st["r"])
    assert np.array_equal(pi, oi) and np.array_equal(pj, oj)

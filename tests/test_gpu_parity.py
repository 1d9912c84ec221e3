"""Parity of the HIP path (through the C ABI) with the oracle and with the reference's golden vectors.

Bars: neighbour pair sets bit-exact; FP64 derivatives, states and time steps within 1e-10 relative
(L2 over the particles, per quantity) of the FP64 CPU result, per step."""
import numpy as np
import pytest

from conftest import rel_l2
from opensph_amd import gpu, initial

pytestmark = pytest.mark.gpu

TOL = 1e-10
DERIV = ("dv", "du", "divv", "drho", "p", "cs", "ds", "gradv", "ddamage", "reduce", "s")
STATE = ("r", "v", "dv", "rho", "drho", "u", "du", "s", "ds", "damage", "ddamage", "divv")


def _cases():
    from cases import exercised_fluid, exercised_impact

    return {
        "fluid10k": lambda: initial.fluid_sphere(),
        "fluid_exercised": lambda: exercised_fluid(),
        "impact_exercised": lambda: exercised_impact(),
    }


def _print_case(name, st):
    print(f"[case] {name} n={st.n}", flush=True)


def _assert_close(a, b, fields, tol=TOL):
    for k in fields:
        x, y = a[k], b[k]
        if k == "dv":
            x, y = x[:, :3], y[:, :3]
        assert rel_l2(x, y) < tol, (k, rel_l2(x, y))


@pytest.mark.parametrize("name", ["fluid10k", "fluid_exercised", "impact_exercised"])
def test_pairs_bit_exact(oracle, name):
    st, cfg = _cases()[name]()
    _print_case(name, st)
    g = gpu.GpuSolver(st, cfg)
    pi, pj = g.pairs()
    oi, oj = oracle.pairs(st["r"])
    assert np.array_equal(pi, oi) and np.array_equal(pj, oj)


@pytest.mark.parametrize("divisor", [1, 2, 3, 4])
def test_pairs_mixed_smoothing_lengths(oracle, divisor):
    """Four smoothing-length classes (h spans a factor 16): every class grid and cross-class search, with
    neighbour cells of 2 h_max / divisor."""
    rng = np.random.default_rng(5)
    st, cfg = initial.fluid_sphere(n_requested=3000, radius=1000.0)
    st["r"][:, :3] += rng.normal(0, 10.0, size=(st.n, 3))
    st["r"][:, 3] *= np.exp2(rng.uniform(-2.0, 2.0, size=st.n))
    _print_case(f"mixed_h/{divisor}", st)
    g = gpu.GpuSolver(st, cfg)
    g.tune(cell_divisor=divisor)
    pi, pj = g.pairs()
    oi, oj = oracle.pairs(st["r"])
    assert len(pi) > 0
    assert np.array_equal(pi, oi) and np.array_equal(pj, oj)


def test_pairs_dense_cluster_grows_lists(oracle):
    """More than the default list capacity of neighbours: the build must grow its lists and stay exact."""
    rng = np.random.default_rng(9)
    n = 700
    st, cfg = initial.fluid_sphere(n_requested=n, radius=10.0)
    st["r"][:, :3] = rng.uniform(-1.0, 1.0, size=(st.n, 3))
    st["r"][:, 3] = 0.8
    _print_case("dense", st)
    g = gpu.GpuSolver(st, cfg)
    pi, pj = g.pairs()
    oi, oj = oracle.pairs(st["r"])
    assert len(pi) / st.n > 100
    assert np.array_equal(pi, oi) and np.array_equal(pj, oj)


def test_pairs_after_grid_tables_grow(oracle):
    """Searches on one context whose grid grows between them: first more cells with the same rows (the x
    extent stretched past a power of two of cells), then more rows.  The cell and row tables are
    reallocated independently; every pair set stays exact."""
    st, cfg = initial.fluid_sphere(n_requested=2000, radius=500.0)
    g = gpu.GpuSolver(st, cfg)
    for stretch in ((1.0, 1.0), (6.0, 1.0), (6.0, 5.0)):
        s2 = st.copy()
        s2["r"][:, 0] *= stretch[0]
        s2["r"][:, 1] *= stretch[1]
        _print_case(f"grid_growth x{stretch[0]} y{stretch[1]}", s2)
        g.upload(s2)
        pi, pj = g.pairs()
        oi, oj = oracle.pairs(s2["r"])
        assert len(pi) > 0
        assert np.array_equal(pi, oi) and np.array_equal(pj, oj)


@pytest.mark.parametrize("name", ["fluid10k", "fluid_exercised", "impact_exercised"])
def test_integrate_matches_oracle(oracle, name):
    st, cfg = _cases()[name]()
    _print_case(name, st)
    g = gpu.GpuSolver(st, cfg)
    g.integrate()
    a = g.download(st)
    b = oracle.integrate(st, cfg)
    _assert_close(a, b, DERIV)
    assert np.array_equal(a["ncnt"], b["ncnt"])


@pytest.mark.parametrize("lanes", [2, 4])
def test_integrate_lane_split_matches_oracle(oracle, lanes):
    """Two and four lanes sharing each particle's neighbour list in the pair kernel (the default is one):
    FP64 derivatives within the parity bar of the oracle, neighbour counts exact."""
    st, cfg = _cases()["impact_exercised"]()
    _print_case(f"impact_exercised/lanes{lanes}", st)
    g = gpu.GpuSolver(st, cfg)
    g.tune(pairs_lanes=lanes)
    g.integrate()
    a = g.download(st)
    b = oracle.integrate(st, cfg)
    _assert_close(a, b, DERIV)
    assert np.array_equal(a["ncnt"], b["ncnt"])


@pytest.mark.parametrize("name,steps", [("fluid10k", 20), ("impact_exercised", 6)])
def test_predictor_corrector_matches_oracle(oracle, name, steps):
    st, cfg = _cases()[name]()
    _print_case(name, st)
    g = gpu.GpuSolver(st, cfg)
    glog = [g.step() for _ in range(steps)]
    so, olog = oracle.run_pc(st, cfg, steps)
    for (dt, nxt, crit), (odt, onxt, ocrit) in zip(glog, olog):
        assert nxt == pytest.approx(onxt, rel=TOL)
        assert crit == ocrit
    _assert_close(g.download(st), so, STATE)


@pytest.mark.parametrize("precision,block", [("fp64", 256), ("mixed", 256), ("fp64", 128), ("mixed", 64)])
def test_fused_finalize_equals_separate_pass(precision, block):
    """The PC step with the finalize, corrector and criteria fused into the pair kernel gives the same bits
    as with them in their own pass (same arithmetic per particle, same order), for every pair-kernel block
    size (the fused criteria minimum is reduced over the block's own waves)."""
    st, cfg = _cases()["impact_exercised"]()
    _print_case(f"fused_finalize/{precision}/block{block}", st)
    out = []
    for fuse in (1, 0):
        g = gpu.GpuSolver(st, cfg, precision=precision)
        g.tune(fuse_finalize=fuse, pairs_block=block if fuse else 256)
        log = [g.step() for _ in range(4)]
        out.append((log, g.download(st)))
        g.close()
    (la, a), (lb, b) = out
    assert la == lb
    for k in STATE:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("name", ["fluid.npz", "impact.npz"])
def test_matches_reference_golden(name):
    from test_oracle import _state_from, golden_case

    g, meta, cfg, st = golden_case(name)
    _print_case(name, st)
    solver = gpu.GpuSolver(st, cfg)
    solver.integrate()
    a = solver.download(st)
    ref = _state_from(g, "integ", meta["materials"])
    _assert_close(a, ref, DERIV)
    assert np.array_equal(a["ncnt"], ref["ncnt"])
    solver.upload(st)
    solver.dt = cfg.initial_dt
    log = [solver.step() for _ in range(int(meta["steps"]))]
    for (dt, nxt, crit), (rdt, rnxt, rcrit) in zip(log, g["dts"]):
        assert nxt == pytest.approx(rnxt, rel=TOL)
        assert crit == int(rcrit)
    _assert_close(solver.download(st), _state_from(g, "steps", meta["materials"]),
                  ("r", "v", "rho", "u", "s", "damage", "dv", "drho", "du", "ds"))


def test_empty_and_single_particle():
    st, cfg = initial.fluid_sphere(n_requested=200, radius=100.0)
    one = st.subset([0])
    _print_case("single", one)
    g = gpu.GpuSolver(one, cfg)
    g.integrate()
    a = g.download(one)
    assert a["ncnt"][0] == 0 and np.all(a["dv"] == 0.0)
    dt, nxt, crit = g.step()
    assert np.isfinite(nxt)
    empty = st.subset([])
    _print_case("empty", empty)
    g = gpu.GpuSolver(empty, cfg)
    g.integrate()
    assert g.pairs()[0].size == 0


@pytest.mark.parametrize("screen", [0, 1])
def test_pair_cutoff_boundary(oracle, screen):
    """Two particles exactly at |r| = h_i + h_j are not neighbours; one ulp closer they are (also
    through the FP32-screened list kernel, where both cases fall inside its error band)."""
    st, cfg = initial.fluid_sphere(n_requested=200, radius=100.0)
    two = st.subset([0, 1])
    two["r"][:] = [[0.0, 0.0, 0.0, 1.0], [2.0, 0.0, 0.0, 1.0]]
    _print_case(f"boundary/screen={screen}", two)
    g = gpu.GpuSolver(two, cfg)
    g.tune(lists_screen=screen)
    assert g.pairs()[0].size == 0
    two["r"][1, 0] = np.nextafter(2.0, 0.0)
    g.upload(two)
    pi, pj = g.pairs()
    assert list(zip(pi, pj)) == [(0, 1)]
    assert oracle.pairs(two["r"])[0].tolist() == [0]


@pytest.mark.parametrize("origin", [0.0, 3.0e4])
def test_pairs_lattice_at_cutoff(oracle, origin):
    """A cubic lattice of spacing 1 with h = 1/2 puts every nearest-neighbour pair exactly at the cutoff;
    perturbations of a few ulps make about half of them neighbours.  Far from the origin the FP32 screen
    band widens with the domain-to-h ratio.  Both list kernels must give the oracle's set."""
    rng = np.random.default_rng(31)
    m = 22
    g1 = np.arange(m, dtype=np.float64)
    xyz = np.stack(np.meshgrid(g1, g1, g1, indexing="ij"), axis=-1).reshape(-1, 3) + origin
    st, cfg = initial.fluid_sphere(n_requested=200, radius=100.0)
    lat = st.subset(np.zeros(len(xyz), dtype=np.int64))
    ulps = rng.integers(-3, 4, size=xyz.shape)
    lat["r"][:, :3] = xyz + ulps * np.spacing(np.maximum(np.abs(xyz), 1.0))
    lat["r"][:, 3] = 0.5
    oi, oj = oracle.pairs(lat["r"])
    assert 0.2 * 3 * len(xyz) < len(oi) < 0.8 * 3 * len(xyz)
    for screen in (0, 1):
        _print_case(f"lattice_cutoff/{origin}/screen={screen}", lat)
        g = gpu.GpuSolver(lat, cfg)
        g.tune(lists_screen=screen)
        pi, pj = g.pairs()
        assert np.array_equal(pi, oi) and np.array_equal(pj, oj), screen


@pytest.mark.slow
def test_full_size_impact_exact_pairs_and_derivatives(oracle):
    """The 1,062,568-particle impact, perturbed so that every term is active: bit-exact pair set and
    FP64 derivatives within 1e-10 of the oracle; momentum conserved to rounding."""
    from cases import exercised_impact

    st, cfg = exercised_impact(n_target=950_000, n_projectile=50_000, seed=17)
    _print_case("impact_1M", st)
    g = gpu.GpuSolver(st, cfg)
    pi, pj = g.pairs()
    oi, oj = oracle.pairs(st["r"])
    assert np.array_equal(pi, oi) and np.array_equal(pj, oj)
    g.tune(lists_screen=1)
    si, sj = g.pairs()
    assert np.array_equal(si, oi) and np.array_equal(sj, oj)
    g.tune(lists_screen=-1)
    g.upload(st)
    g.integrate()
    a = g.download(st)
    b = oracle.integrate(st, cfg)
    _assert_close(a, b, DERIV)
    assert np.array_equal(a["ncnt"], b["ncnt"])
    p = (a["m"][:, None] * a["dv"][:, :3]).sum(axis=0)
    scale = (a["m"][:, None] * np.abs(a["dv"][:, :3])).sum(axis=0)
    assert np.all(np.abs(p) <= 1e-11 * scale)


@pytest.mark.slow
def test_full_size_impact_pc_steps_match_oracle(oracle):
    """Two FP64 PredictorCorrector steps of the 1,062,568-particle impact, perturbed so that every term and
    criterion is active, against the oracle's steps: time steps and criteria equal, the state and its
    rates within 1e-10 (the damage at the parity bar except where a flaw sits on its activation
    threshold, at most one particle in a thousand; see test_gpu_configs.py)."""
    from cases import exercised_impact

    st, cfg = exercised_impact(n_target=950_000, n_projectile=50_000, seed=17)
    _print_case("impact_1M_pc", st)
    g = gpu.GpuSolver(st, cfg)
    log = [g.step() for _ in range(2)]
    a = g.download(st)
    b, rlog = oracle.run_pc(st, cfg, 2)
    for (dt, nx, cr), (odt, onx, ocr) in zip(log, rlog):
        assert dt == pytest.approx(odt, rel=1e-12) and nx == pytest.approx(onx, rel=1e-10) and cr == ocr
    for k in ("r", "v", "dv", "rho", "drho", "u", "du", "s", "ds"):
        x, y = (a[k][:, :3], b[k][:, :3]) if k in ("v", "dv") else (a[k], b[k])
        assert rel_l2(x, y) <= 1e-10, (k, rel_l2(x, y))
    off = np.abs(a["damage"] - b["damage"]) > 1e-10 * max(float(np.abs(b["damage"]).max()), 1e-300)
    assert off.sum() <= st.n // 1000
    assert rel_l2(a["damage"][~off], b["damage"][~off]) <= 1e-10
    g.close()


@pytest.mark.parametrize("fuse,gravity", [(1, 0), (0, 0), (1, 1)])
def test_pc_steps_through_list_overflow_match_oracle(oracle, fuse, gravity):
    """PC steps whose neighbour lists overflow inside the step (more neighbours than the lists hold): the
    step grows the lists and repeats its second half from the predicted state, with the self-gravity
    copied in again when it is attached.  Overflows on the first step (capacity 96, ~150 neighbours) and
    again after a denser state is uploaded into the same context; every step matches the oracle."""
    import dataclasses

    st, cfg = initial.fluid_sphere(n_requested=3000, radius=1000.0, sigma_v=10.0)
    st["r"][:, 3] *= 1.7
    cfg = dataclasses.replace(cfg, gravity=gravity)
    _print_case(f"list overflow in the step fuse={fuse} gravity={gravity}", st)
    g = gpu.GpuSolver(st, cfg)
    g.tune(fuse_finalize=fuse)
    for k, s in enumerate((st, None)):
        if s is None:
            s = st.copy()
            s["r"][:, 3] *= 1.4  # ~2.7x the neighbours: the grown lists overflow again
            g.upload(s)
            g.dt = cfg.initial_dt
        glog = [g.step() for _ in range(3)]
        so, olog = oracle.run_pc(s, cfg, 3)
        for (dt, nxt, crit), (odt, onxt, ocrit) in zip(glog, olog):
            assert nxt == pytest.approx(onxt, rel=TOL), k
            assert crit == ocrit, k
        out = g.download(s)
        _assert_close(out, so, STATE)
        assert np.array_equal(out["ncnt"], so["ncnt"])
        assert out["ncnt"].max() > (96 if k == 0 else 200)
    g.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_list_overflow_inside_the_step_is_bit_exact(precision):
    """A PC step whose lists overflow and are grown inside the step gives the same bits as the same step
    on a context whose lists were grown beforehand (by a neighbour search)."""
    st, cfg = initial.fluid_sphere(n_requested=3000, radius=1000.0, sigma_v=10.0)
    st["r"][:, 3] *= 1.7
    _print_case(f"list overflow bit-exact {precision}", st)
    a = gpu.GpuSolver(st, cfg, precision=precision)
    b = gpu.GpuSolver(st, cfg, precision=precision)
    b.pairs()  # grows b's lists before its first step
    la = [a.step() for _ in range(3)]
    lb = [b.step() for _ in range(3)]
    assert la == lb
    da, db = a.download(st), b.download(st)
    assert da["ncnt"].max() > 96
    for k in STATE + ("ncnt", "p", "cs"):
        assert np.array_equal(da[k], db[k]), k
    a.close()
    b.close()


@pytest.mark.slow
def test_full_size_impact_steps():
    """The unperturbed 1M impact: neighbour statistics of the reference probe (29 / 66.3 / 68), first
    step leaves the bodies undeformed (reference Impact test, core/sph/solvers/test/Impact.cpp:41-58),
    and a few PC steps stay finite."""
    st, cfg = initial.impact()
    _print_case("impact_1M_plain", st)
    g = gpu.GpuSolver(st, cfg)
    stats = g.integrate()
    assert stats["neighbors_min"] == 29 and stats["neighbors_max"] == 68
    assert stats["neighbors_mean"] == pytest.approx(66.289, abs=1e-3)
    g.upload(st)
    g.step()
    a = g.download(st)
    assert np.all(a["gradv"] == 0.0) and np.all(a["s"] == 0.0) and np.all(a["rho"] == 2700.0)
    for _ in range(3):
        g.step()
    a = g.download(st)
    for k in ("r", "v", "rho", "u", "s"):
        assert np.all(np.isfinite(a[k])), k


def test_integrals_match_state_and_conserve_momentum():
    """osph_integrals (core/physics/Integrals.cpp:24-180) against sums over the downloaded state, and the
    symmetric pair forces conserve the total momentum over predictor-corrector steps."""
    from cases import exercised_impact

    st, cfg = exercised_impact()
    _print_case("integrals", st)
    g = gpu.GpuSolver(st, cfg)
    for _ in range(3):
        g.step()
    a = g.download(st)
    it = g.integrals()
    m, v, r = a["m"], a["v"][:, :3], a["r"][:, :3]
    ke = 0.5 * np.sum(m * np.sum(v * v, axis=1))
    ie = np.sum(m * a["u"])
    assert it["kinetic_energy"] == pytest.approx(ke, rel=1e-12)
    assert it["internal_energy"] == pytest.approx(ie, rel=1e-12)
    assert it["energy"] == pytest.approx(ke + ie, rel=1e-12)
    scale = np.sum(m[:, None] * np.abs(v), axis=0)
    assert np.all(np.abs(it["momentum"] - np.sum(m[:, None] * v, axis=0)) <= 1e-12 * scale)
    L = np.sum(m[:, None] * np.cross(r, v), axis=0)
    assert np.allclose(it["angular_momentum"], L, rtol=1e-10, atol=1e-12 * np.abs(L).max())
    assert np.allclose(it["center_of_mass"], np.sum(m[:, None] * r, axis=0) / m.sum(), rtol=1e-12, atol=1e-6)
    assert it["mass"] == pytest.approx(m.sum(), rel=1e-14)
    p0 = g.integrals()["momentum"]
    for _ in range(3):
        g.step()
    p1 = g.integrals()["momentum"]
    assert np.all(np.abs(p1 - p0) <= 1e-10 * scale)

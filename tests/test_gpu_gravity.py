"""The HIP self-gravity against the gravity oracle, through the C ABI (include/osph_gravity.h): the direct
sum, Barnes-Hut at theta -> 0 (the direct sum exactly), Barnes-Hut within the tolerances of the
reference's own tests (core/gravity/test/BarnesHut.cpp), root moments, point evaluation, energy,
LeapFrog and integrals, edge cases (0, 1, 2 particles, coincident positions)."""
import numpy as np
import pytest

from oracle.bridge import GravityOracle
from opensph_amd.gravity import G_SI, GpuGravity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grav():
    return GravityOracle()


def ball(n, radius, seed, h=None):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, size=(4 * n + 64, 3))
    x = x[np.linalg.norm(x, axis=1) < 1][:n] * radius
    hh = h if h is not None else radius * (4.0 / 3.0 * np.pi / n) ** (1 / 3)
    return np.c_[x, np.full(len(x), hh)]


def gas_storage(n=1000):
    """As getGravityStorage (BarnesHut.cpp test :14-20): r0 = 1e7 m, rho0 = 100, equal masses."""
    r = ball(n, 1e7, 31)
    m = np.full(len(r), 100.0 * 4.0 / 3.0 * np.pi * 1e21 / len(r))
    return r, m


def almost_equal(a, b, eps):
    """tests/Approx.h semantics for vectors: |a - b|^2 <= eps^2 (1 + max(|a|^2, |b|^2))."""
    d2 = np.sum((a - b) ** 2, axis=1)
    return d2 <= eps ** 2 * (1.0 + np.maximum(np.sum(a ** 2, axis=1), np.sum(b ** 2, axis=1)))


def case(name, **kw):
    print(f"[case] gravity {name} {kw}", flush=True)


@pytest.mark.parametrize("kernel", ["point", "m4"])
def test_direct_sum_matches_oracle(grav, kernel):
    r, m = gas_storage(2000)
    r[:, 3] = 4e5
    case("direct", kernel=kernel, n=len(m))
    g = GpuGravity(r, m, solver="brute_force", kernel=kernel)
    g.build()
    a, _ = g.eval_self()
    ref = grav.direct(r, m, G_SI, kernel=kernel)
    assert np.max(np.abs(a - ref)) <= 1e-14 * np.max(np.abs(ref))
    sub = np.array([0, 17, 999, len(m) - 1], dtype=np.uint32)
    assert np.max(np.abs(g.direct(sub) - ref[sub])) <= 1e-14 * np.max(np.abs(ref))


@pytest.mark.parametrize("order", [0, 2, 3])
def test_barnes_hut_zero_opening_angle_is_direct_sum(grav, order):
    """'BarnesHut zero opening angle' (BarnesHut.cpp test :47-82): theta = EPS opens every node."""
    r, m = gas_storage(1500)
    case("bh_theta0", order=order)
    g = GpuGravity(r, m, theta=1e-12, order=order, leaf_size=5, kernel="point")
    g.build()
    a, st = g.eval_self()
    ref = grav.direct(r, m, G_SI, kernel="point")
    assert st["nodes_approx"] == 0
    assert np.max(np.abs(a - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("order,eps", [(0, 3e-2), (2, 3e-3), (3, 3e-3)])
def test_barnes_hut_ellipsoid_within_reference_tolerance(grav, order, eps):
    """'BarnesHut storage acceleration' (BarnesHut.cpp test :162-213): theta 0.4, leaf 5, an ellipsoid."""
    r, m = gas_storage(1000)
    r[:, :3] *= [2.0, 0.5, 0.1]
    case("bh_ellipsoid", order=order)
    g = GpuGravity(r, m, theta=0.4, order=order, leaf_size=5, kernel="point")
    g.build()
    a, st = g.eval_self()
    ref = grav.direct(r, m, G_SI, kernel="point")
    assert st["nodes_approx"] > 0
    assert not np.any(np.all(a == ref, axis=1))
    assert np.all(almost_equal(a, ref, eps))


def test_barnes_hut_error_shrinks_with_opening_angle_and_order(grav):
    r, m = gas_storage(3000)
    ref = grav.direct(r, m, G_SI)
    err = {}
    for theta in (0.8, 0.4, 0.2):
        for order in (0, 2, 3):
            case("bh_convergence", theta=theta, order=order)
            g = GpuGravity(r, m, theta=theta, order=order, leaf_size=10)
            g.build()
            a, _ = g.eval_self()
            err[theta, order] = np.sqrt(np.mean(np.sum((a - ref) ** 2, axis=1)) / np.mean(np.sum(ref ** 2, axis=1)))
    for order in (0, 2, 3):
        assert err[0.2, order] < err[0.4, order] < err[0.8, order], err
    for theta in (0.8, 0.4, 0.2):
        assert err[theta, 3] < err[theta, 2] < err[theta, 0], err


def test_root_moments_match_direct_moments(grav):
    """'BarnesHut storage moments' (BarnesHut.cpp test :110-121): root moments of the tree against the
    moments of all particles computed directly, 1e-10."""
    r, m = gas_storage(1000)
    case("moments")
    g = GpuGravity(r, m, theta=0.5, order=3, leaf_size=5)
    g.build()
    got = g.moments()
    want = grav.moments(r, m)
    scale = np.abs(want[4:16]).max()
    assert got[0] == pytest.approx(want[0], rel=1e-10)  # masses are premultiplied by G, then divided
    assert np.allclose(got[1:4], want[1:4], rtol=1e-12, atol=1e-12 * 1e7)
    assert np.max(np.abs(got[4:16] - want[4:16])) <= 1e-10 * scale


@pytest.mark.parametrize("order", [0, 2, 3])
def test_point_evaluation_known_answer(order):
    """'BarnesHut simple acceleration' (BarnesHut.cpp test :123-141): six unit masses, leaf 1,
    evalAcceleration at (-10, 10, 0) / G.  The expected value is the exact sum; at theta -> 0 every
    leaf is summed and it is reproduced to rounding.  At theta = 0.5 the reference's tolerances (4e-4,
    8e-5, 1e-5) are those of its k-d tree's nodes for this one point; the radix tree approximates other
    nodes, so the bound here is the truncation scale of the criterion size / distance < theta: a
    relative error below theta^(p+2) / 2 after order p (the dipole term vanishes): 12.5 %, 3.1 %, 1.6 %."""
    r = np.array([[2, 3, 0, 1], [5, 4, 0, 1], [9, 6, 0, 1], [4, 7, 0, 1], [8, 1, 0, 1], [7, 2, 0, 1]], float)
    expected = np.array([0.020169998934707004, -0.007912678499211458, 0.0])
    case("kat", order=order)
    g = GpuGravity(r, np.ones(6), theta=1e-12, order=order, leaf_size=1, kernel="point")
    g.build()
    a = g.eval_points(np.array([[-10.0, 10.0, 0.0]]))[0] / G_SI
    assert np.allclose(a, expected, rtol=0, atol=1e-16)
    g = GpuGravity(r, np.ones(6), theta=0.5, order=order, leaf_size=1, kernel="point")
    g.build()
    a = g.eval_points(np.array([[-10.0, 10.0, 0.0]]))[0] / G_SI
    assert not np.array_equal(a, expected)
    bound = 0.5 ** ({0: 0, 2: 2, 3: 3}[order] + 2) / 2
    assert np.linalg.norm(a - expected) < bound * np.linalg.norm(expected), (a, expected)


@pytest.mark.parametrize("order", [0, 2, 3])
def test_far_field_point_is_the_root_expansion(grav, order):
    """Far from a compact cluster the whole tree is one accepted node: evalAcceleration returns the
    expansion of the root moments (evaluateGravity, Moments.h:318-360), which must equal the oracle's
    expansion of the directly computed moments."""
    rng = np.random.default_rng(61)
    r = np.c_[rng.uniform(-1, 1, (60, 3)), np.full(60, 0.1)]
    m = rng.uniform(0.5, 1.5, 60)
    case("far_field", order=order)
    g = GpuGravity(r, m, theta=0.5, order=order, leaf_size=4, kernel="point", G=1.0)
    g.build()
    mo = grav.moments(r, m)
    for p in ([40.0, -25.0, 30.0], [-60.0, 10.0, -5.0], [3.0, 80.0, 1.0]):
        got = g.eval_points(np.array([p]))[0]
        want = grav.multipole(np.r_[mo[0], mo[4:16]], np.array(p) - mo[1:4], order)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-14 * np.linalg.norm(want)), (p, got, want)


def test_point_evaluation_matches_self_gravity():
    """'BarnesHut eval evalAll equality' (BarnesHut.cpp test :215-238): evalAcceleration at each particle
    agrees with evalSelfGravity within 1e-2 (different opening criteria)."""
    r, m = gas_storage(1000)
    case("points_vs_self")
    g = GpuGravity(r, m, theta=0.25, order=3, leaf_size=10, kernel="m4")
    g.build()
    a, _ = g.eval_self()
    # evalAcceleration at a particle's position includes the particle itself, whose softened
    # contribution vanishes at zero separation
    b = g.eval_points(r[:, :3])
    assert np.all(almost_equal(a, b, 1e-2))


def test_symmetrisation_and_accumulation():
    """'BarnesHut symmetrization' and 'override accelerations bug' (BarnesHut.cpp test :315-356)."""
    r = np.array([[0.0, 0.0, 0.0, 1.0], [2.0, 0.0, 0.0, 5.0]])
    case("symmetrisation")
    g = GpuGravity(r, np.full(2, 1e10), theta=0.5, order=3)
    g.build()
    a, _ = g.eval_self()
    assert np.array_equal(a[0], -a[1]) and a[0, 0] > 0
    g = GpuGravity(r, np.full(2, 1e-12), theta=0.5, order=3)
    g.build()
    dv = np.array([[3.0, 1.0, 1.0], [4.0, -2.0, 10.0]])
    out, _ = g.eval_self(dv)
    assert np.array_equal(out, dv)


def test_edge_cases(grav):
    case("edge")
    g = GpuGravity(np.zeros((0, 4)), np.zeros(0), theta=0.5)
    g.build()
    a, _ = g.eval_self()
    assert a.shape == (0, 3)
    one = np.array([[1.0, 2.0, 3.0, 0.5]])
    g = GpuGravity(one, np.ones(1), theta=0.5)
    g.build()
    assert np.array_equal(g.eval_self()[0], np.zeros((1, 3)))
    assert np.allclose(g.moments()[:4], [1.0, 1.0, 2.0, 3.0])
    # a cluster of coincident particles (equal Morton codes) plus a distant group
    rng = np.random.default_rng(3)
    r = np.r_[np.tile([[0.0, 0.0, 0.0, 1.0]], (300, 1)), np.c_[rng.uniform(50, 60, (200, 3)), np.ones(200)]]
    m = rng.uniform(1, 2, len(r))
    g = GpuGravity(r, m, theta=1e-12, order=2, leaf_size=8, G=1.0)
    g.build()
    a, _ = g.eval_self()
    ref = grav.direct(r, m, 1.0)
    assert np.max(np.abs(a - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_energy_and_integrals(grav):
    r, m = gas_storage(1200)
    rng = np.random.default_rng(5)
    v = rng.normal(0, 10.0, (len(m), 3))
    case("energy")
    g = GpuGravity(r, m, v=v, solver="brute_force")
    e = g.energy()
    assert e == pytest.approx(grav.energy(r, m, G_SI), rel=1e-13)
    it = g.integrals()
    assert it["kinetic_energy"] == pytest.approx(0.5 * np.sum(m * np.sum(v ** 2, axis=1)), rel=1e-13)
    assert np.allclose(it["momentum"], (m[:, None] * v).sum(0), rtol=1e-12, atol=1e-9 * np.abs(m[:, None] * v).sum())
    L = (m[:, None] * np.cross(r[:, :3], v)).sum(0)
    assert np.allclose(it["angular_momentum"], L, rtol=1e-12, atol=1e-12 * np.abs(L).max())
    assert np.allclose(it["center_of_mass"], (m[:, None] * r[:, :3]).sum(0) / m.sum(), rtol=1e-12, atol=1e-3)


@pytest.mark.parametrize("solver", ["brute_force", "barnes_hut"])
def test_leapfrog_matches_oracle(grav, solver):
    """LeapFrog + HardSphereSolver gravity (TimeStepping.cpp:352-391, NBodySolver.cpp:97-117): five
    steps against the oracle's LeapFrog with the direct sum (Barnes-Hut at theta -> 0)."""
    r = ball(800, 1e5, 41, h=2e3)
    rng = np.random.default_rng(42)
    v = np.c_[rng.normal(0, 20.0, (len(r), 3)), np.zeros(len(r))]
    m = np.full(len(r), 1e15)
    case("leapfrog", solver=solver)
    g = GpuGravity(r, m, v=v, solver=solver, theta=1e-12, order=2, leaf_size=25)
    g.leapfrog(1.0, 5)
    r1, v1, a1 = grav.leapfrog(r, v, m, G_SI, 1.0, 5)
    assert np.allclose(g.get("r")[:, :3], r1[:, :3], rtol=0, atol=1e-9 * 1e5)
    assert np.allclose(g.get("v")[:, :3], v1[:, :3], rtol=0, atol=1e-11 * np.abs(v1).max())
    assert np.allclose(g.get("dv"), a1, rtol=0, atol=1e-11 * np.abs(a1).max())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_leapfrog_in_morton_order_gives_the_same_bits(precision):
    """LeapFrog steps with the resident state moved into the tree's Morton order at every tree build give the
    same positions, velocities and accelerations (returned in the caller's order) as steps that never move
    it; the calls that take or return per-particle data in the caller's order (direct sums on sampled
    targets, accelerations added to a given dv, a new set of particles) go through the permutation."""
    r = ball(5000, 1e5, 61, h=10.0)
    rng = np.random.default_rng(62)
    v = np.c_[rng.normal(0, 30.0, (len(r), 3)), np.zeros(len(r))]
    m = rng.uniform(0.5e12, 2e12, len(r))
    case("leapfrog morton order", precision=precision)
    runs = []
    for period in (1, 0):
        g = GpuGravity(r, m, v=v, theta=0.5, order=2, leaf_size=25, precision=precision)
        g.tune(reorder_period=period)
        g.leapfrog(1.0, 6)
        runs.append(g)
    a, b = runs
    for q in ("r", "v", "dv"):
        assert np.array_equal(a.get(q), b.get(q)), q
    sample = rng.choice(len(r), 300, replace=False).astype(np.uint32)
    da, db = a.direct(sample), b.direct(sample)
    assert np.allclose(da, db, rtol=1e-12, atol=1e-12 * np.abs(db).max())
    dv0 = rng.normal(0, 1e-3, (len(r), 3))
    ea, _ = a.eval_self(dv0)
    eb, _ = b.eval_self(dv0)
    assert np.array_equal(ea, eb)
    assert a.energy() == pytest.approx(b.energy(), rel=1e-12)
    ia, ib = a.integrals(), b.integrals()
    assert ia["kinetic_energy"] == pytest.approx(ib["kinetic_energy"], rel=1e-12)
    # new particles after a move: taken and returned in the caller's order again
    a.set_particles(r, m, v=v)
    assert np.array_equal(a.get("r")[:, :3], r[:, :3])
    assert np.array_equal(a.get("v")[:, :3], v[:, :3])


def test_large_barnes_hut_against_direct_sample():
    """A 300k-particle uniform sphere (the shape of the 8M N-body workload at reduced size): quadrupole
    Barnes-Hut at theta 0.5, leaf 25, M4 softening, against the direct sum on 512 sampled targets."""
    n = 300_000
    r = ball(n, 1e5, 51, h=10.0)
    m = np.full(len(r), 1.25e12)
    case("large", n=len(r))
    g = GpuGravity(r, m, theta=0.5, order=2, leaf_size=25)
    g.build()
    a, st = g.eval_self()
    sample = np.random.default_rng(52).choice(len(r), 512, replace=False).astype(np.uint32)
    d = g.direct(sample)
    rel = np.linalg.norm(a[sample] - d, axis=1) / np.linalg.norm(d, axis=1)
    assert np.median(rel) < 2e-3 and rel.max() < 3e-2, (np.median(rel), rel.max())
    assert st["nodes_approx"] > 0 and st["nodes_exact"] > 0

"""The gravity oracle (oracle/osph_oracle_gravity.c) pinned without a GPU: against the reference's own
known answers (core/gravity/test/BarnesHut.cpp), closed forms of the softened kernel, the convergence
order of the multipole expansion, invariants of the direct sum, and (where it was built from source)
the real reference's BruteForceGravity and BarnesHut."""
import numpy as np
import pytest

from oracle.bridge import GravityOracle, RefRun, ref_gravity

G = 6.67408e-11


@pytest.fixture(scope="module")
def grav():
    return GravityOracle()


def six_masses():
    """getTestParticles (core/gravity/test/BarnesHut.cpp:23-45)"""
    r = np.array([[2, 3, 0, 1], [5, 4, 0, 1], [9, 6, 0, 1], [4, 7, 0, 1], [8, 1, 0, 1], [7, 2, 0, 1]], float)
    return r, np.ones(6)


def ball(n, radius, seed, h=None):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, size=(4 * n, 3))
    x = x[np.linalg.norm(x, axis=1) < 1][:n] * radius
    hh = h if h is not None else radius * (4.0 / 3.0 * np.pi / n) ** (1 / 3)
    return np.c_[x, np.full(len(x), hh)]


def test_point_kat_is_the_direct_sum(grav):
    """The expected acceleration of the reference's 'simple acceleration' test is the exact direct sum."""
    r, m = six_masses()
    a = grav.points(r, m, 1.0, np.array([[-10.0, 10.0, 0.0, 1.0]]), kernel="point")[0]
    assert np.allclose(a, [0.020169998934707004, -0.007912678499211458, 0.0], rtol=0, atol=1e-17)


def test_softened_kernel_closed_forms(grav):
    """GravityKernel<CubicSpline<3>> is the potential of the M4 density: psi(2) = -1/2, psi'(2) = 1/4
    (Newtonian outside), grad(0) = 4/3 = 4 pi W(0)/3, and the Poisson equation
    (1/q^2) d/dq (q^2 psi') = 4 pi W_M4(q) holds across the table."""
    v, g = grav.values, grav.grads
    assert v[-1] == pytest.approx(-0.5, abs=1e-15)
    assert g[-1] == pytest.approx(0.125, abs=1e-15)  # grad = psi'/q = 1/q^3 at q = 2
    assert g[0] == 4.0 / 3.0
    q2 = np.arange(40001) / 10000.0
    q = np.sqrt(q2)
    dpsi = g * q  # psi'(q)
    # finite-difference Poisson check away from the table's ends and the M4 break at q = 1
    idx = np.arange(200, 39800, 37)
    idx = idx[np.abs(q[idx] - 1.0) > 0.02]
    lhs = []
    for i in idx:
        qa, qb = q[i - 1], q[i + 1]
        lhs.append((qb * qb * dpsi[i + 1] - qa * qa * dpsi[i - 1]) / (qb - qa) / (q[i] ** 2))
    w = np.where(q[idx] < 1, 0.25 * (2 - q[idx]) ** 3 - (1 - q[idx]) ** 3, 0.25 * (2 - q[idx]) ** 3) / np.pi
    assert np.allclose(lhs, 4 * np.pi * w, rtol=2e-3, atol=2e-4)


def test_pair_energy_and_force_closed_forms(grav):
    """Two point masses: U = -G m1 m2 / r, a = G m / r^2 towards the other one; with the M4 kernel and
    r < 2 hbar the force is softened below the Newtonian value."""
    r = np.array([[0.0, 0.0, 0.0, 1.0], [3.0, 0.0, 0.0, 1.0]])
    m = np.array([2.0, 5.0])
    assert grav.energy(r, m, 1.0, kernel="point") == pytest.approx(-10.0 / 3.0, rel=1e-15)
    a = grav.direct(r, m, 1.0, kernel="point")
    assert np.allclose(a, [[5.0 / 9.0, 0, 0], [-2.0 / 9.0, 0, 0]], rtol=1e-15)
    near = np.array([[0.0, 0.0, 0.0, 1.0], [1.0, 0.0, 0.0, 1.0]])
    a = grav.direct(near, m, 1.0)
    assert 0.0 < a[0, 0] < 5.0 and a[0, 0] == pytest.approx(5.0 * grav.grads[10000], rel=1e-12)


def test_direct_sum_conserves_momentum(grav):
    r = ball(400, 1e4, 3, h=500.0)
    m = np.random.default_rng(4).uniform(1e10, 2e10, size=len(r))
    a = grav.direct(r, m, G)
    p = (m[:, None] * a).sum(axis=0)
    assert np.all(np.abs(p) <= 1e-13 * (m[:, None] * np.abs(a)).sum(axis=0))


def test_symmetrised_pair_is_antisymmetric(grav):
    """'BarnesHut symmetrization' (BarnesHut.cpp test :315-331): h = 1 and 5, equal masses."""
    r = np.array([[0.0, 0.0, 0.0, 1.0], [2.0, 0.0, 0.0, 5.0]])
    a = grav.direct(r, np.full(2, 1e10), G)
    assert np.array_equal(a[0], -a[1])


def test_moments_closed_forms(grav):
    """Two equal masses at +-s on x: M = 2, com = 0, q2 = diag(4/3, -2/3, -2/3) s^2, q3 = 0.  Masses 1
    and 3 at x = 0 and 1: com = 3/4, the raw M3_xxx = -3/8 and its traceless reduction has
    q3_xxx = 2/5 M3_xxx and q3_xyy = -1/5 M3_xxx (the -1/5 trace factor over three index placements)."""
    s = 3.0
    r = np.array([[-s, 0, 0, 1], [s, 0, 0, 1]], float)
    mo = grav.moments(r, np.ones(2))
    assert mo[0] == 2.0 and np.allclose(mo[1:4], 0.0)
    assert np.allclose(mo[4:9], [4 / 3 * s * s, 0, 0, -2 / 3 * s * s, 0], atol=1e-14)
    assert np.allclose(mo[9:16], 0.0, atol=1e-14)
    mo = grav.moments(np.array([[0, 0, 0, 1], [1, 0, 0, 1]], float), np.array([1.0, 3.0]))
    assert np.allclose(mo[1:4], [0.75, 0, 0])
    m3 = -0.375
    assert np.allclose(mo[9:16], [0.4 * m3, 0, 0, -0.2 * m3, 0, 0, 0], atol=1e-15)


@pytest.mark.parametrize("order,expo", [(0, 2), (2, 3), (3, 4)])
def test_multipole_error_converges_with_its_order(grav, order, expo):
    """For a cluster of size s seen from distance d, the expansion truncated after order p errs by
    O((s/d)^(p+1)) relative to the direct sum (the dipole vanishes about the centre of mass, so the
    monopole error is already second order).  A dropped or mis-scaled term breaks the slope."""
    r = ball(60, 1.0, 11)
    m = np.random.default_rng(12).uniform(0.5, 1.5, len(r))
    mo = grav.moments(r, m)
    mom13 = np.r_[mo[0], mo[4:16]]
    com = mo[1:4]
    direction = np.array([0.6, -0.48, 0.64])
    errs = []
    for d in (8.0, 16.0, 32.0, 64.0):
        p = com + d * direction
        exact = grav.points(r, m, 1.0, np.array([[*p, 0.0]]), kernel="point")[0]
        approx = grav.multipole(mom13, p - com, order)
        errs.append(np.linalg.norm(approx - exact) / np.linalg.norm(exact))
    slopes = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(slopes > expo - 0.35), (errs, slopes)


def test_leapfrog_is_time_reversible(grav):
    """Drift-kick-drift is symmetric: dt steps followed by -dt steps return to the start."""
    r = ball(40, 1e3, 5, h=50.0)
    rng = np.random.default_rng(6)
    v = np.c_[rng.normal(0, 1e-3, (len(r), 3)), np.zeros(len(r))]
    m = np.full(len(r), 1e12)
    r1, v1, _ = grav.leapfrog(r, v, m, G, 2.0, 10)
    r2, v2, _ = grav.leapfrog(r1, -v1, m, G, 2.0, 10)
    assert np.allclose(r2[:, :3], r[:, :3], rtol=0, atol=1e-9 * 1e3)


def test_leapfrog_circular_orbit():
    """Two equal point masses on a circular orbit: after one period they are back within O(dt^2)."""
    grav = GravityOracle()
    m = np.array([1.0, 1.0])
    sep = 2.0
    omega = np.sqrt(2.0 / sep ** 3)  # G (m1 + m2) / sep^3
    vel = omega * sep / 2
    r = np.array([[-1.0, 0, 0, 1e-3], [1.0, 0, 0, 1e-3]])
    v = np.array([[0, -vel, 0, 0], [0, vel, 0, 0]])
    period = 2 * np.pi / omega
    steps = 2000
    r1, v1, _ = grav.leapfrog(r, v, m, 1.0, period / steps, steps)
    assert np.allclose(r1[:, :3], r[:, :3], atol=1e-4)


needs_ref = pytest.mark.skipif(not RefRun.available(), reason="reference core not built (oracle/_ref)")


@needs_ref
@pytest.mark.parametrize("kernel", ["point", "m4"])
def test_direct_sum_equals_reference_bruteforce(grav, kernel):
    r = ball(300, 1e7, 21, h=8e5)
    m = np.full(len(r), 1e19)
    a_ref, _ = ref_gravity(r, m, G, solver="brute_force", kernel=kernel)
    assert np.array_equal(grav.direct(r, m, G, kernel=kernel), a_ref)


@needs_ref
def test_reference_barnes_hut_tolerances_against_direct_sum(grav):
    """The reference BarnesHut within the tolerances of its own 'storage acceleration' test (theta = 0.4,
    leaf 5, an ellipsoid), measured against this oracle's direct sum; at theta -> 0 it is the direct sum."""
    r = ball(1000, 1e7, 22, h=1e5)
    r[:, :3] *= [2.0, 0.5, 0.1]
    m = np.full(len(r), 100.0 * 4 / 3 * np.pi * 1e21 / len(r))
    direct = grav.direct(r, m, G, kernel="point")
    for order, eps in ((0, 3e-2), (2, 3e-3), (3, 3e-3)):
        a, _ = ref_gravity(r, m, G, kernel="point", theta=0.4, order=order, leaf=5)
        diff2 = np.sum((a - direct) ** 2, axis=1)
        bound = eps ** 2 * (1 + np.maximum(np.sum(a ** 2, axis=1), np.sum(direct ** 2, axis=1)))
        assert np.all(diff2 <= bound), order
    a0, mom = ref_gravity(r, m, G, kernel="point", theta=1e-12, order=3, leaf=5)
    assert np.allclose(a0, direct, rtol=1e-12, atol=1e-15 * np.abs(direct).max())
    om = grav.moments(r, m)
    assert np.allclose(mom, np.r_[om[0], om[4:16]], rtol=1e-10, atol=1e-10 * np.abs(mom).max())

"""Initial conditions for the hot-path workloads (setup only; never on the timed path).

The generators follow the reference's setup rules so that particle counts and layouts match what
OpenSPH would build for the same request:

* hexagonal close packing inside a sphere: HexagonalPacking::generate
  (core/sph/initial/Distribution.cpp:126-201), lattice loop of Box::iterateWithIndices
  (core/objects/geometry/Box.h:182-195), centring on the domain centre;
* smoothing length times eta and masses proportional to h^3 normalised to rho0 * V:
  InitialConditions::setQuantities / getMasses (core/sph/initial/Initial.cpp:293-333);
* two-body collision placement: CollisionGeometrySetupJob::evaluate
  (core/run/jobs/ParticleJobs.cpp:532-577), bounding spheres from getBoundingSphere (:499-510);
* Grady-Kipp flaws, accumulate mode: ScalarGradyKippModel::setFlaws (core/physics/Damage.cpp:17-126),
  with this module's own seeded generator in place of the reference RNG;
* minimal density of ContinuityEquation::create (core/sph/equations/EquationTerm.cpp:318-336).

Random inputs (velocity perturbation, flaw draws) come from numpy's PCG64 with an explicit seed; the
same arrays are fed to every implementation that is compared.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import math
from typing import List, Tuple

import numpy as np

from .state import INFTY, ParticleState, RunConfig, basalt

PI = 3.14159265358979323846264338327950288  # core/math/MathUtils.h:379


def sphere_volume(radius: float) -> float:
    """sphereVolume (core/math/MathUtils.h:393-395)."""
    return 1.3333333333333333333333 * PI * (radius * radius * radius)


_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.cbrt.restype = ctypes.c_double
_libm.cbrt.argtypes = [ctypes.c_double]


def cbrt(x: float) -> float:
    """root<3> (core/math/MathUtils.h:104-107): the C library's cbrt, which numpy's vectorised np.cbrt does
    not always reproduce to the last bit."""
    return float(_libm.cbrt(float(x)))


def _axis_values(lo: float, hi: float, step: float) -> List[float]:
    out = []
    v = lo
    while v <= hi:
        out.append(v)
        v += step
    return out


def hcp_sphere(n_requested: int, radius: float, center=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Hexagonal close packing inside a sphere; returns (N, 4) positions with lane 3 = h (before eta)."""
    volume = sphere_volume(radius)
    particle_density = float(n_requested) / volume
    h = 1.0 / cbrt(particle_density)
    dx = 1.1 * h
    dy = math.sqrt(3.0) * 0.5 * dx
    dz = math.sqrt(6.0) / 3.0 * dx
    c = np.asarray(center, dtype=np.float64)
    lower = c - radius
    upper = c + radius
    lo = lower + 0.5 * np.array([dx, dy, dz])
    xs = np.array(_axis_values(lo[0], upper[0], dx))
    ys = np.array(_axis_values(lo[1], upper[1], dy))
    zs = np.array(_axis_values(lo[2], upper[2], dz))
    delta_x = 0.5 * dx
    delta_y = math.sqrt(3.0) / 6.0 * dx
    K, J, I = np.meshgrid(np.arange(len(zs)), np.arange(len(ys)), np.arange(len(xs)), indexing="ij")
    X = xs[I].copy()
    Y = ys[J].copy()
    Z = zs[K].copy()
    even_k = (K % 2) == 0
    shift_x = np.where(even_k, (J % 2) == 1, (J % 2) == 0)
    X[shift_x] += delta_x
    Y[~even_k] += delta_y
    X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
    ddx, ddy, ddz = X - c[0], Y - c[1], Z - c[2]
    inside = (ddx * ddx + ddy * ddy + ddz * ddz) <= radius * radius
    pos = np.stack([X[inside], Y[inside], Z[inside]], axis=1)
    # centre of mass of the lattice onto the domain centre (sequential sum, as the reference loop)
    com = np.array([np.cumsum(pos[:, k])[-1] for k in range(3)]) / pos.shape[0]
    pos = pos + (c - com)
    out = np.empty((pos.shape[0], 4))
    out[:, :3] = pos
    out[:, 3] = h
    return out


def _masses(h: np.ndarray, total_mass: float) -> np.ndarray:
    m = h * h * h
    prelim = np.cumsum(m)[-1]
    return m * (total_mass / prelim)


def sphere_body(n_requested: int, radius: float, material: dict, center=(0.0, 0.0, 0.0), eta: float = 1.3):
    """Positions (N,4) with h *= eta and masses for one monolithic spherical body."""
    r = hcp_sphere(n_requested, radius, center)
    r[:, 3] *= eta
    m = _masses(r[:, 3], sphere_volume(radius) * material["rho0"])
    return r, m


def young_modulus(material: dict) -> float:
    """Young modulus set by ScalarGradyKippModel::setFlaws (core/physics/Damage.cpp:43-48)."""
    mu, A = material["shear_modulus"], material["A"]
    return mu * 9.0 * A / (3.0 * A + mu)


def grady_kipp_flaws(r: np.ndarray, m: np.ndarray, rho: np.ndarray, material: dict, rng: np.random.Generator,
                     k_weibull: float = 4.0e35, m_weibull: float = 9.0, rayleigh: float = 0.4,
                     kernel_radius: float = 2.0):
    """Flaw activation thresholds, Weibull accumulate mode (core/physics/Damage.cpp:17-126).

    Returns (eps_min, m_zero, growth, n_flaws)."""
    n = r.shape[0]
    mu, A, rho0 = material["shear_modulus"], material["A"], material["rho0"]
    cg = rayleigh * math.sqrt((A + 4.0 / 3.0 * mu) / rho0)
    growth = cg / (kernel_radius * r[:, 3])
    V = np.cumsum(m / rho)[-1]
    denom = 1.0 / (math.pow(k_weibull, 1.0 / m_weibull) * math.pow(V, 1.0 / m_weibull))
    # draw particle indices until every particle holds at least one flaw
    chunks = []
    covered = np.zeros(n, dtype=bool)
    n_cov = 0
    total = 0
    while n_cov < n:
        draw = (rng.random(max(4 * n, 1 << 16)) * n).astype(np.int64)
        first_new = []
        chunks.append(draw)
        for start in range(0, draw.size, 1 << 20):
            part = draw[start:start + (1 << 20)]
            uniq, first = np.unique(part, return_index=True)
            newmask = ~covered[uniq]
            covered[uniq[newmask]] = True
            n_cov += int(newmask.sum())
            if n_cov >= n:
                # the loop stops right after the draw that flaws the last particle
                last_pos = start + int(first[newmask].max())
                chunks[-1] = draw[: last_pos + 1]
                break
        total += chunks[-1].size
    idx = np.concatenate(chunks)
    K = idx.size
    order = np.argsort(idx, kind="stable")
    sidx = idx[order]
    bounds = np.flatnonzero(np.diff(sidx)) + 1
    starts = np.concatenate([[0], bounds])
    ends = np.concatenate([bounds, [K]]) - 1
    part = sidx[starts]
    first_p = order[starts] + 1  # flaw number p is 1-based
    last_p = order[ends] + 1
    n_flaws = np.zeros(n, dtype=np.uint32)
    n_flaws[part] = (ends - starts + 1).astype(np.uint32)
    eps_min = np.zeros(n)
    eps_max = np.zeros(n)
    eps_min[part] = denom * np.power(first_p.astype(np.float64), 1.0 / m_weibull)
    eps_max[part] = denom * np.power(last_p.astype(np.float64), 1.0 / m_weibull)
    m_zero = np.ones(n)
    many = n_flaws > 1
    m_zero[many] = np.log(n_flaws[many].astype(np.float64)) / np.log(eps_max[many] / eps_min[many])
    return eps_min, m_zero, growth, n_flaws


def bodies(st: ParticleState) -> List[Tuple[int, int]]:
    """(first, count) of each body: the runs of equal FLAG, as the bodies are laid out in storage."""
    flag = np.asarray(st["flag"])
    if flag.size == 0:
        return []
    cuts = np.concatenate([[0], np.flatnonzero(np.diff(flag)) + 1, [flag.size]])
    return [(int(a), int(b - a)) for a, b in zip(cuts[:-1], cuts[1:])]


def device_flaws(solver, st: ParticleState, seed: int, k_weibull: float = 4.0e35, m_weibull: float = 9.0,
                 rayleigh: float = 0.4):
    """The Grady-Kipp flaws of every body of `st` (resident in `solver`) drawn on the device
    (osph_set_flaws): each body's volume sum(m / rho) and crack-growth speed as grady_kipp_flaws takes
    them; body b draws with seed + b."""
    for b, (first, count) in enumerate(bodies(st)):
        sl = slice(first, first + count)
        material = st.materials[int(st["mat"][first])]
        mu, A, rho0 = material["shear_modulus"], material["A"], material["rho0"]
        cg = rayleigh * math.sqrt((A + 4.0 / 3.0 * mu) / rho0)
        V = float(np.cumsum(st["m"][sl] / st["rho"][sl])[-1])
        solver.set_flaws(first, count, V, cg, seed + b, k_weibull=k_weibull, m_weibull=m_weibull)


def contact_time(st: ParticleState, kernel_radius: float = 2.0) -> float:
    """Earliest time at which a particle of one body comes within the kernel support of a particle of
    another, with every body moving rigidly at its particles' mean velocity (the state of a collision set
    up by CollisionGeometrySetupJob, core/run/jobs/ParticleJobs.cpp:532-577, before the bodies touch).
    The support of a pair is kernel_radius * hbar = (h_i + h_j) for the M4 kernel; the smallest h of each
    body is used, so at the returned time at least one pair is inside its support.  Until then every pair
    of the run joins particles of one body, and a body at rest in its reference state (rho = rho0, u = 0,
    S = 0) has zero pair sums.  Returns 0 for fewer than two bodies, inf when the bodies do not meet."""
    from scipy.spatial import cKDTree

    parts = bodies(st)
    if len(parts) < 2:
        return 0.0
    r = np.asarray(st["r"])
    v = np.asarray(st["v"])
    best = math.inf
    for a in range(len(parts)):
        for b in range(a + 1, len(parts)):
            fa, ca = parts[a]
            fb, cb = parts[b]
            pa, pb = r[fa:fa + ca, :3], r[fb:fb + cb, :3]
            w = v[fb:fb + cb, :3].mean(axis=0) - v[fa:fa + ca, :3].mean(axis=0)
            cut = 0.5 * kernel_radius * (float(r[fa:fa + ca, 3].min()) + float(r[fb:fb + cb, 3].min()))
            tree = cKDTree(pa)

            def gap(t):
                d, _ = tree.query(pb + w * t, k=1)
                return float(d.min()) - cut

            g0 = gap(0.0)
            if g0 < 0.0:
                return 0.0
            speed = float(np.linalg.norm(w))
            if speed == 0.0:
                continue
            # the gap closes at most at the relative speed, so a step of gap / speed never passes the
            # contact (sphere tracing); the floor of 1 % of the support bounds the step count, and the
            # contact is then bisected inside the last step
            reach = g0 + float(np.ptp(pa, axis=0).max()) + float(np.ptp(pb, axis=0).max())
            lo, hi, g = 0.0, math.inf, g0
            for _ in range(10000):
                t = lo + max(g, 0.01 * cut) / speed
                g = gap(t)
                if g < 0.0:
                    hi = t
                    break
                if g > reach:
                    break
                lo = t
            if not math.isfinite(hi):
                continue
            for _ in range(50):
                mid = 0.5 * (lo + hi)
                if gap(mid) < 0.0:
                    hi = mid
                else:
                    lo = mid
            best = min(best, hi)
    return best


def kernel_w0() -> float:
    """Central value of the M4 LUT kernel, LutKernel::valueImpl(0) = CubicSpline<3>::valueImpl(0)."""
    q = 0.0
    return (1.0 / PI) * (0.25 * (2.0 - q) ** 3 - (1.0 - q) ** 3)


def density_lower_bound(r: np.ndarray, m: np.ndarray, material: dict) -> float:
    """ContinuityEquation::create: max(min_i m_i w0 / h_i^3, DENSITY_RANGE.lower)."""
    w0 = kernel_w0()
    h = r[:, 3]
    limit = float(np.min(m * w0 / (h * h * h)))
    return max(limit, material["rho_lo"])


def _fill_body(st: ParticleState, sl: slice, r, m, flag: int, mat: int, material: dict):
    st["r"][sl] = r
    st["m"][sl] = m
    st["flag"][sl] = flag
    st["mat"][sl] = mat
    st["rho"][sl] = material["rho0"]
    st["u"][sl] = 0.0
    st["reduce"][sl] = 1.0


def fluid_sphere(n_requested: int = 10000, radius: float = 1000.0, sigma_v: float = 1.0, seed: int = 1,
                 material: dict | None = None) -> Tuple[ParticleState, RunConfig]:
    """Config: one basalt sphere, PressureForce + StandardAV + ContinuityEquation, ConstSmoothingLength,
    Courant criterion only; zero internal energy and a seeded Gaussian velocity perturbation."""
    mat = basalt(**(material or {}))
    r, m = sphere_body(n_requested, radius, mat)
    st = ParticleState(r.shape[0], [mat])
    _fill_body(st, slice(None), r, m, 0, 0, mat)
    rng = np.random.default_rng(seed)
    st["v"][:, :3] = rng.normal(0.0, sigma_v, size=(st.n, 3))
    lo = density_lower_bound(st["r"], st["m"], mat)
    st.materials[0]["rho_lo"] = lo
    cfg = RunConfig(solid_stress=0, adaptive_h=0, has_reduce=0, has_damage=0, criteria=1)
    return st, cfg


def gas_collision(n_total: int = 500_000, radius: float = 1000.0, ratio: float = 0.5, speed: float = 100.0,
                  u0: float = 1.0e3, gap_h: float = 0.5, sigma_v: float = 1.0,
                  seed: int = 5) -> Tuple[ParticleState, RunConfig]:
    """Config (the adjoint run): two basalt spheres of radii R and ratio R at equal particle spacing,
    PressureForce + StandardAV + ContinuityEquation, Tillotson, no stress, damage or plasticity
    (EosMaterial), ConstSmoothingLength, Courant criterion.  The smaller body approaches the larger one
    along -x at `speed`, its surface `gap_h` smoothing lengths from the other's.  Both start at specific
    internal energy u0 > 0, away from the energy range's lower bound 0, where the clamp of the integrator
    has no derivative.  A seeded Gaussian velocity perturbation (sigma_v per component) keeps the pairs
    off the StandardAV switch (v_i - v_j).(r_i - r_j) = 0, where a rigid body's pairs would all sit and
    the run has only one-sided derivatives."""
    mat = basalt()
    n_t = max(1, int(round(n_total / (1.0 + ratio ** 3))))
    n_p = max(1, n_total - n_t)
    rt, mt = sphere_body(n_t, radius, mat)
    rp, mp = sphere_body(n_p, radius * ratio, mat)
    h = rt[0, 3]
    # place the smaller body on +x, its surface gap_h h away from the larger body's surface
    ext_t = float(np.max(np.sqrt(np.sum(rt[:, :3] ** 2, axis=1))))
    ext_p = float(np.max(np.sqrt(np.sum(rp[:, :3] ** 2, axis=1))))
    rp[:, 0] += ext_t + ext_p + gap_h * h
    nt, npj = rt.shape[0], rp.shape[0]
    st = ParticleState(nt + npj, [mat])
    _fill_body(st, slice(0, nt), rt, mt, 0, 0, mat)
    _fill_body(st, slice(nt, nt + npj), rp, mp, 1, 0, mat)
    st["v"][nt:, 0] = -speed
    if sigma_v > 0.0:
        st["v"][:, :3] += np.random.default_rng(seed).normal(0.0, sigma_v, size=(st.n, 3))
    st["u"][:] = u0
    st.materials[0]["rho_lo"] = density_lower_bound(st["r"], st["m"], mat)
    cfg = RunConfig(solid_stress=0, adaptive_h=0, has_reduce=0, has_damage=0, criteria=1)
    return st, cfg


def impact(n_target: int = 950_000, n_projectile: int = 50_000, r_target: float = 50e3, r_projectile: float = 5e3,
           speed: float = 5e3, angle_deg: float = 45.0, offset_h: float = 4.0, seed: int = 1234,
           flaws: bool = True) -> Tuple[ParticleState, RunConfig]:
    """Config: two-body solid impact, SolidStressForce + StandardAV + ContinuityEquation, Hooke + von Mises +
    scalar Grady-Kipp damage, AdaptiveSmoothingLength, all time-step criteria."""
    mat = basalt(yielding=1, damage=1)
    mat["young_modulus"] = young_modulus(mat)
    rt, mt = sphere_body(n_target, r_target, mat)
    rp, mp = sphere_body(n_projectile, r_projectile, mat)

    def bounding(r, m):
        # getCenterOfMass then the largest distance (ParticleJobs.cpp:499-510)
        c = np.array([np.cumsum(m * r[:, k])[-1] for k in range(3)]) / np.cumsum(m)[-1]
        d = np.sqrt(np.sum((r[:, :3] - c) ** 2, axis=1))
        return c, float(d.max())

    ct, rad_t = bounding(rt, mt)
    cp, rad_p = bounding(rp, mp)
    rt[:, :3] -= ct
    dist = rad_t + rad_p
    h = rt[0, 3]
    phi = angle_deg * (PI / 180.0)
    x = dist * math.cos(phi) + offset_h * h
    y = dist * math.sin(phi)
    rp[:, :3] += -cp + np.array([x, y, 0.0])

    nt, npj = rt.shape[0], rp.shape[0]
    st = ParticleState(nt + npj, [mat, dict(mat)])
    _fill_body(st, slice(0, nt), rt, mt, 0, 0, mat)
    _fill_body(st, slice(nt, nt + npj), rp, mp, 1, 1, mat)
    st["v"][nt:, 0] = -speed
    lo = density_lower_bound(st["r"], st["m"], mat)  # over the whole storage, as solver.create runs on it
    for mm in st.materials:
        mm["rho_lo"] = lo
    rng = np.random.default_rng(seed)
    for sl in (slice(0, nt), slice(nt, nt + npj)):
        if flaws:
            e, mz, g, nf = grady_kipp_flaws(st["r"][sl], st["m"][sl], st["rho"][sl], mat, rng)
        else:
            e, mz, g, nf = (np.full(sl.stop - sl.start, 1.0), np.ones(sl.stop - sl.start),
                            np.zeros(sl.stop - sl.start), np.ones(sl.stop - sl.start, np.uint32))
        st["eps_min"][sl] = e
        st["m_zero"][sl] = mz
        st["growth"][sl] = g
        st["n_flaws"][sl] = nf
    cfg = RunConfig(solid_stress=1, adaptive_h=1, has_reduce=1, has_damage=1, criteria=15)
    return st, cfg


def nbody_sphere(n: int, radius: float = 1.0e5, mass: float = 1.25e12, h: float = 10.0, seed: int = 7):
    """Self-gravity N-body input: n equal masses uniform in a sphere (rejection sampling of numpy's PCG64
    with an explicit seed), at rest, smoothing length h (the M4 softening support is 2 h).  Returns
    (r (n, 4) with h in lane 3, m (n,))."""
    rng = np.random.default_rng(seed)
    out = np.empty((0, 3))
    while len(out) < n:
        x = rng.uniform(-1.0, 1.0, size=(int(1.3 * (n - len(out))) + 1024, 3))
        out = np.concatenate([out, x[np.einsum("ij,ij->i", x, x) < 1.0]])
    r = np.c_[out[:n] * radius, np.full(n, h)]
    return r, np.full(n, mass)


def ensemble_specs(count: int, seed: int = 11, n_range=(5e3, 2e5), offset_h: float = 4.0) -> List[dict]:
    """The members of the impact ensemble as impact() arguments: N ~ exp(U(ln n_lo, ln n_hi)),
    R_p / R_t ~ U(0.05, 0.2) at equal particle spacing (N_p / N_t = (R_p / R_t)^3, N_p >= 100),
    angle ~ U(0, 90) deg, speed ~ U(3, 7) km/s, R_t = 50 km, basalt; one seed per member."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = float(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        ratio = rng.uniform(0.05, 0.2)
        n_p = max(100, int(n * ratio ** 3 / (1 + ratio ** 3)))
        n_t = max(100, int(n) - n_p)
        out.append(dict(n_target=n_t, n_projectile=n_p, r_target=50e3, r_projectile=50e3 * ratio,
                        speed=rng.uniform(3e3, 7e3), angle_deg=rng.uniform(0.0, 90.0),
                        seed=int(rng.integers(1 << 30)), offset_h=offset_h))
    return out


def impact_ensemble(count: int, seed: int = 11, n_range=(5e3, 2e5), flaws: bool = True, offset_h: float = 4.0):
    """Impact ensemble (ensemble_specs) built on the host.  flaws=False leaves the Grady-Kipp flaws to the
    device (device_flaws after the upload).  Returns a list of (state, config)."""
    return [impact(flaws=flaws, **spec) for spec in ensemble_specs(count, seed, n_range, offset_h)]


def impact_setup(n_target: int = 950_000, n_projectile: int = 50_000, r_target: float = 50e3,
                 r_projectile: float = 5e3, speed: float = 5e3, angle_deg: float = 45.0, offset_h: float = 4.0,
                 seed: int = 1234):
    """impact()'s configuration as the arguments of osph_setup_spheres: (bodies, materials, config,
    collision).  The device builds the same particles as impact() bit for bit."""
    mat = basalt(yielding=1, damage=1)
    mat["young_modulus"] = young_modulus(mat)
    bodies = [dict(particle_count=n_target, radius=r_target, eta=1.3, material=0, flag=0),
              dict(particle_count=n_projectile, radius=r_projectile, eta=1.3, velocity=(-speed, 0.0, 0.0),
                   material=1, flag=1)]
    cfg = RunConfig(solid_stress=1, adaptive_h=1, has_reduce=1, has_damage=1, criteria=15)
    return bodies, [mat, dict(mat)], cfg, (angle_deg, offset_h)


def device_impact(precision: str = "fp64", flaw_seed: int = 0, **kw):
    """impact(**kw) built on the device (osph_setup_spheres) and resident in a new GpuSolver, the Grady-Kipp
    flaws of body b drawn on the device with seed flaw_seed + b (as device_flaws draws them).  Returns the
    solver."""
    from .gpu import GpuSolver

    bodies, mats, cfg, col = impact_setup(**kw)
    solver, info = GpuSolver.from_bodies(bodies, mats, cfg, precision=precision, collision=col)
    for b, (first, count, volume, _) in enumerate(info):
        material = solver.materials[b]
        mu, A, rho0 = material["shear_modulus"], material["A"], material["rho0"]
        cg = 0.4 * math.sqrt((A + 4.0 / 3.0 * mu) / rho0)
        solver.set_flaws(first, count, volume, cg, flaw_seed + b)
    return solver

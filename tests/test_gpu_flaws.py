"""Grady-Kipp flaws drawn on the device (osph_set_flaws; ScalarGradyKippModel::setFlaws,
core/physics/Damage.cpp:17-126, accumulate mode): against a numpy restatement of the same counter-based
draws (n_flaws exact, thresholds to libm rounding), independent of the particle order in HBM, the
stopping rule (every particle flawed, the last draw flaws the last one), and the same statistics as the
host generator's flaws."""
import math

import numpy as np
import pytest

from opensph_amd import gpu, initial

pytestmark = pytest.mark.gpu

GOLD = np.uint64(0x9E3779B97F4A7C15)
M32 = np.uint64(0xFFFFFFFF)


def flaw_target(seed, p, nb):
    """splitmix64 of seed + p * golden increment, times nb, high word (numpy restatement)."""
    with np.errstate(over="ignore"):
        x = np.uint64(seed) + p.astype(np.uint64) * GOLD
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
        n = np.uint64(nb)
        hi = (x >> np.uint64(32)) * n + (((x & M32) * n) >> np.uint64(32))
    return (hi >> np.uint64(32)).astype(np.int64)


def restated(nb, seed, volume, k_weibull=4.0e35, m_weibull=9.0):
    first = np.full(nb, np.iinfo(np.int64).max)
    p0 = 1
    while np.any(first == np.iinfo(np.int64).max):
        p = np.arange(p0, p0 + 4 * nb, dtype=np.int64)
        np.minimum.at(first, flaw_target(seed, p, nb), p)
        p0 += 4 * nb
    K = int(first.max())
    p = np.arange(1, K + 1, dtype=np.int64)
    idx = flaw_target(seed, p, nb)
    cnt = np.bincount(idx, minlength=nb)
    last = np.zeros(nb, dtype=np.int64)
    np.maximum.at(last, idx, p)
    denom = 1.0 / (math.pow(k_weibull, 1.0 / m_weibull) * math.pow(volume, 1.0 / m_weibull))
    emin = denom * np.power(first.astype(np.float64), 1.0 / m_weibull)
    emax = denom * np.power(last.astype(np.float64), 1.0 / m_weibull)
    mz = np.ones(nb)
    many = cnt > 1
    mz[many] = np.log(cnt[many].astype(np.float64)) / np.log(emax[many] / emin[many])
    return emin, mz, cnt, K


def _state(n_target=3000, n_projectile=300):
    return initial.impact(n_target=n_target, n_projectile=n_projectile, seed=3, flaws=False)


def test_device_flaws_match_restatement():
    st, cfg = _state()
    print(f"[case] device flaws n={st.n}", flush=True)
    g = gpu.GpuSolver(st, cfg)
    initial.device_flaws(g, st, seed=77)
    out = g.download(st)
    g.close()
    for b, (first, count) in enumerate(initial.bodies(st)):
        sl = slice(first, first + count)
        V = float(np.cumsum(st["m"][sl] / st["rho"][sl])[-1])
        emin, mz, cnt, K = restated(count, 77 + b, V)
        assert np.array_equal(out["n_flaws"][sl], cnt.astype(np.uint32))
        assert int(cnt.sum()) == K and K >= count
        assert np.allclose(out["eps_min"][sl], emin, rtol=1e-14, atol=0.0)
        assert np.allclose(out["m_zero"][sl], mz, rtol=1e-12, atol=0.0)
        mat = st.materials[0]
        cg = 0.4 * math.sqrt((mat["A"] + 4.0 / 3.0 * mat["shear_modulus"]) / mat["rho0"])
        assert np.allclose(out["growth"][sl], cg / (2.0 * st["r"][sl, 3]), rtol=1e-15, atol=0.0)


def test_device_flaws_independent_of_particle_order():
    """After steps that reorder the resident state, the flaws written to each particle are the same."""
    st, cfg = _state()
    print(f"[case] device flaws reorder n={st.n}", flush=True)
    a = gpu.GpuSolver(st, cfg)
    a.tune(reorder_period=1)
    for _ in range(3):
        a.step()
    moved = a.download(st)
    initial.device_flaws(a, st, seed=5)
    x = a.download(st)
    b = gpu.GpuSolver(moved, cfg)
    initial.device_flaws(b, st, seed=5)
    y = b.download(st)
    for k in ("eps_min", "m_zero", "growth", "n_flaws"):
        assert np.array_equal(x[k], y[k]), k
    a.close()
    b.close()


def test_device_flaws_statistics_match_host_generator():
    """A 200k-particle body, device and host generators each against the law of the draws: the flaw count
    K is the coupon collector's (K/n - ln n is Gumbel: mean 0.5772, sd pi/sqrt 6), given K the counts
    per particle have the binomial variance K/n (1 - 1/n), and a particle's first flaw over n is Exp(1)
    (median ln 2), which eps_min = denom first^(1/m) carries."""
    st, cfg = initial.impact(n_target=200_000, n_projectile=2000, seed=9, flaws=True)
    print(f"[case] device flaws statistics n={st.n}", flush=True)
    g = gpu.GpuSolver(st, cfg)
    initial.device_flaws(g, st, seed=11)
    dev = g.download(st)
    g.close()
    first, count = initial.bodies(st)[0]
    sl = slice(first, first + count)
    V = float(np.cumsum(st["m"][sl] / st["rho"][sl])[-1])
    denom = 1.0 / (math.pow(4.0e35, 1.0 / 9.0) * math.pow(V, 1.0 / 9.0))
    for name, out in (("device", dev), ("host", st)):
        nf = out["n_flaws"][sl].astype(np.float64)
        assert np.all(nf >= 1), name
        lam = nf.sum() / count
        assert abs(lam - math.log(count) - 0.5772) < 5 * math.pi / math.sqrt(6.0), (name, lam)
        assert abs(nf.var() - lam * (1 - 1 / count)) < 0.05 * lam, (name, nf.var(), lam)
        first_flaw = (out["eps_min"][sl] / denom) ** 9.0
        assert abs(np.median(first_flaw) / count - math.log(2.0)) < 0.02, (name, np.median(first_flaw) / count)


def test_set_flaws_arguments():
    st, cfg = _state(400, 40)
    print(f"[case] device flaws arguments n={st.n}", flush=True)
    g = gpu.GpuSolver(st, cfg)
    before = g.download(st)
    g.set_flaws(0, 0, 1.0, 1.0, seed=1)  # an empty body changes nothing
    after = g.download(st)
    for k in ("eps_min", "m_zero", "growth", "n_flaws"):
        assert np.array_equal(after[k], before[k]), k
    for args in ((st.n, 1, 1.0, 1.0), (0, st.n + 1, 1.0, 1.0), (0, 10, 0.0, 1.0), (0, 10, 1.0, -1.0)):
        with pytest.raises(gpu.GpuError):
            g.set_flaws(*args, seed=1)
    g.set_flaws(0, 1, 1.0, 1.0, seed=1)  # a one-particle body: one flaw, m_zero 1
    one = g.download(st)
    assert one["n_flaws"][0] == 1 and one["m_zero"][0] == 1.0
    g.close()

"""The FP32 Barnes-Hut evaluation (osph_gravity_set_precision(OSPH_PRECISION_MIXED)) against the FP64 one
on the same tree and interaction lists, and against the oracle's direct sum where every node is opened:
multipole orders 0, 2 and 3, point and M4-softened kernels, systems from metres to astronomical units
(lengths and G m are scaled by powers of two), a clustered system, the 300k sphere, and LeapFrog steps."""
import numpy as np
import pytest

from oracle.bridge import GravityOracle
from opensph_amd.gravity import G_SI, GpuError, GpuGravity

pytestmark = pytest.mark.gpu

# FP32 terms summed in batches of up to 64 into FP64, close pairs separated in FP64: the acceleration
# error, measured against the largest acceleration of the system, is a few FP32 ulps (2e-7 measured on
# a 2M-particle sphere)
TOL = 1e-6


def ball(n, radius, seed, h=None):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, size=(4 * n + 64, 3))
    x = x[np.linalg.norm(x, axis=1) < 1][:n] * radius
    hh = h if h is not None else radius * (4.0 / 3.0 * np.pi / n) ** (1 / 3)
    return np.c_[x, np.full(len(x), hh)]


def both(r, m, **kw):
    out = {}
    for prec in ("fp64", "mixed"):
        g = GpuGravity(r, m, precision=prec, **kw)
        g.build()
        out[prec] = g.eval_self()
        g.close()
    return out


def err(a, ref):
    scale = np.linalg.norm(ref, axis=1).max()
    return np.linalg.norm(a - ref, axis=1).max() / scale


@pytest.mark.parametrize("kernel", ["point", "m4"])
@pytest.mark.parametrize("order", [0, 2, 3])
def test_mixed_matches_fp64_barnes_hut(order, kernel):
    r = ball(3000, 1e7, 31)
    m = np.full(len(r), 100.0 * 4.0 / 3.0 * np.pi * 1e21 / len(r))
    print(f"[case] gravity mixed order={order} kernel={kernel} n={len(r)}", flush=True)
    out = both(r, m, theta=0.5, order=order, leaf_size=10, kernel=kernel)
    (a64, st64), (amx, stmx) = out["fp64"], out["mixed"]
    assert st64 == stmx  # the same tree and interaction lists
    assert st64["nodes_approx"] > 0 and st64["pairs"] > 0
    assert err(amx, a64) < TOL, err(amx, a64)


def test_mixed_with_every_node_opened_matches_direct_sum():
    """theta -> 0: the FP32 evaluation is the direct sum of the oracle to FP32 accuracy."""
    r = ball(1500, 1e7, 33)
    m = np.full(len(r), 1e20)
    print(f"[case] gravity mixed theta0 n={len(r)}", flush=True)
    g = GpuGravity(r, m, theta=1e-12, order=2, leaf_size=5, kernel="point", precision="mixed")
    g.build()
    a, st = g.eval_self()
    assert st["nodes_approx"] == 0
    ref = GravityOracle().direct(r, m, G_SI, kernel="point")
    assert err(a, ref) < TOL, err(a, ref)


@pytest.mark.parametrize("leaf", [33, 47, 64])
def test_large_leaves_with_every_node_opened_match_direct_sum(leaf):
    """Target leaves of more than 32 particles, up to the 64-particle limit, take one split of the
    wavefront's lanes (the others idle); theta -> 0 in both precisions is the oracle's direct sum."""
    r = ball(2000, 1e7, 37)
    m = np.full(len(r), 1e20)
    ref = GravityOracle().direct(r, m, G_SI, kernel="point")
    for prec, tol in (("fp64", 1e-13), ("mixed", TOL)):
        print(f"[case] gravity theta0 leaf={leaf} precision={prec} n={len(r)}", flush=True)
        g = GpuGravity(r, m, theta=1e-12, order=2, leaf_size=leaf, kernel="point", precision=prec)
        g.build()
        a, st = g.eval_self()
        g.close()
        assert st["nodes_approx"] == 0
        assert err(a, ref) < tol, (prec, err(a, ref))


@pytest.mark.parametrize("radius,mass", [(1.0, 1e-3), (1e5, 1e12), (1.5e11, 1e29), (1e17, 1e33)])
def test_mixed_scales_from_metres_to_astronomical(radius, mass):
    """Lengths and G m are brought into FP32's range by powers of two: the same relative accuracy at
    every scale (unscaled, 1/d^5 of the octupole term leaves FP32's range at the larger scales)."""
    r = ball(2000, radius, 35)
    m = np.full(len(r), mass)
    print(f"[case] gravity mixed scale radius={radius} mass={mass}", flush=True)
    out = both(r, m, theta=0.6, order=3, leaf_size=8, kernel="m4")
    a64, amx = out["fp64"][0], out["mixed"][0]
    assert np.all(np.isfinite(amx))
    assert err(amx, a64) < TOL, err(amx, a64)


def test_mixed_clustered_system():
    """Two dense clumps of 1 m far apart (1e5 m) and a diffuse halo: leaves of very different sizes, pairs
    much closer than their offsets from the target leaf (taken in FP64)."""
    rng = np.random.default_rng(36)
    a = ball(1500, 1.0, 37, h=0.05)
    b = ball(1500, 1.0, 38, h=0.05)
    b[:, 0] += 1e5
    halo = ball(1000, 3e5, 39, h=100.0)
    r = np.vstack([a, b, halo])
    m = rng.uniform(0.5, 2.0, size=len(r)) * 1e9
    print(f"[case] gravity mixed clustered n={len(r)}", flush=True)
    out = both(r, m, theta=0.5, order=2, leaf_size=16, kernel="m4")
    a64, amx = out["fp64"][0], out["mixed"][0]
    for part in (slice(0, 1500), slice(1500, 3000), slice(3000, None)):
        assert err(amx[part], a64[part]) < TOL, (part, err(amx[part], a64[part]))


def test_mixed_large_sphere():
    """The 300k-particle sphere of the N-body workload's shape: mixed against FP64 on every particle."""
    r = ball(300_000, 1e5, 51, h=10.0)
    m = np.full(len(r), 1.25e12)
    print(f"[case] gravity mixed large n={len(r)}", flush=True)
    out = both(r, m, theta=0.5, order=2, leaf_size=25, kernel="m4")
    a64, amx = out["fp64"][0], out["mixed"][0]
    assert err(amx, a64) < TOL, err(amx, a64)
    rel = np.linalg.norm(amx - a64) / np.linalg.norm(a64)
    assert rel < 2e-7, rel


def test_mixed_leapfrog_follows_fp64():
    r = ball(2000, 1e5, 41, h=2e3)
    rng = np.random.default_rng(42)
    v = np.c_[rng.normal(0, 20.0, (len(r), 3)), np.zeros(len(r))]
    m = np.full(len(r), 1e15)
    print(f"[case] gravity mixed leapfrog n={len(r)}", flush=True)
    runs = {}
    for prec in ("fp64", "mixed"):
        g = GpuGravity(r, m, v=v, theta=0.5, order=2, leaf_size=25, precision=prec)
        g.leapfrog(1.0, 10)
        runs[prec] = (g.get("r")[:, :3], g.get("v")[:, :3])
        g.close()
    (r64, v64), (rmx, vmx) = runs["fp64"], runs["mixed"]
    dr = np.abs(rmx - (r[:, :3])).max()
    assert np.abs(rmx - r64).max() <= 1e-5 * dr
    assert np.abs(vmx - v64).max() <= 1e-5 * np.abs(v64 - v[:, :3]).max()


def test_set_precision_rejects_unknown_value():
    r = ball(100, 1.0, 43)
    g = GpuGravity(r, np.ones(len(r)))
    with pytest.raises(GpuError):
        g._check(g.L.osph_gravity_set_precision(g.ctx, 7), "osph_gravity_set_precision")
    g.close()

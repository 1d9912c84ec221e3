"""The FP32-screened Barnes-Hut walk (osph_gravity_tune "walk_screen") produces the interaction lists of
the FP64 walk: with the same (FP64) evaluation, the accelerations are bit-identical and the interaction
counts equal, over opening angles, leaf sizes, orders, clustered and coincident particles, far-off-origin
systems and scales from metres to 1e17 m."""
import numpy as np
import pytest

from opensph_amd.gravity import GpuError, GpuGravity

pytestmark = pytest.mark.gpu


def ball(n, radius, seed, h=None):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, size=(4 * n + 64, 3))
    x = x[np.linalg.norm(x, axis=1) < 1][:n] * radius
    hh = h if h is not None else radius * (4.0 / 3.0 * np.pi / n) ** (1 / 3)
    return np.c_[x, np.full(len(x), hh)]


def walks_agree(r, m, precision="fp64", **kw):
    out = {}
    for screen in (0, 1):
        g = GpuGravity(r, m, precision=precision, **kw)
        g.tune(walk_screen=screen)
        g.build()
        out[screen] = g.eval_self()
        g.close()
    (a0, s0), (a1, s1) = out[0], out[1]
    assert s0 == s1, (s0, s1)
    assert np.array_equal(a0, a1)
    return s0


@pytest.mark.parametrize("theta,leaf,order", [(0.5, 25, 2), (0.3, 5, 3), (0.8, 10, 0), (1e-12, 5, 2), (1.5, 64, 2)])
def test_screened_walk_equals_fp64_walk(theta, leaf, order):
    r = ball(20000, 1e5, 61, h=10.0)
    m = np.full(len(r), 1.25e12)
    print(f"[case] walk theta={theta} leaf={leaf} order={order}", flush=True)
    st = walks_agree(r, m, theta=theta, leaf_size=leaf, order=order)
    assert st["pairs"] > 0


@pytest.mark.parametrize("radius,offset", [(1.0, 0.0), (1e5, 1e7), (1.5e11, 0.0), (1e17, 3e17)])
def test_screened_walk_scales_and_offsets(radius, offset):
    """Systems far from the origin (the frame is the root box's centre) and at every scale."""
    r = ball(8000, radius, 62)
    r[:, 0] += offset
    m = np.full(len(r), 1.0)
    print(f"[case] walk radius={radius} offset={offset}", flush=True)
    st = walks_agree(r, m, theta=0.5, leaf_size=16, order=2)
    assert st["nodes_approx"] > 0


def test_screened_walk_clustered_and_coincident():
    """Dense clumps, a halo, a lattice (many ties between box faces and opening balls) and more than
    leaf_size coincident particles (r_open = 0 nodes)."""
    a = ball(3000, 1.0, 63, h=0.05)
    b = ball(3000, 1.0, 64, h=0.05)
    b[:, 0] += 1e5
    halo = ball(2000, 3e5, 65, h=100.0)
    g = np.arange(12, dtype=float)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * 1000.0 + 2e5
    lat = np.c_[lat, np.full(len(lat), 400.0)]
    same = np.tile([[5e4, 5e4, 5e4, 10.0]], (40, 1))
    r = np.vstack([a, b, halo, lat, same])
    m = np.random.default_rng(66).uniform(0.5, 2.0, size=len(r)) * 1e9
    print(f"[case] walk clustered n={len(r)}", flush=True)
    for theta in (0.5, 1.0):
        walks_agree(r, m, theta=theta, leaf_size=8, order=2, kernel="m4")


def test_screened_walk_in_mixed_precision_is_the_default():
    r = ball(20000, 1e5, 67, h=10.0)
    m = np.full(len(r), 1.25e12)
    print("[case] walk mixed default", flush=True)
    walks_agree(r, m, precision="mixed", theta=0.5, leaf_size=25, order=2)
    g0 = GpuGravity(r, m, precision="mixed")
    g0.build()
    a0, _ = g0.eval_self()
    g0.tune(walk_screen=1)
    a1, _ = g0.eval_self()
    assert np.array_equal(a0, a1)
    with pytest.raises(GpuError):
        g0.tune(walk_screen=2)
    g0.close()

"""osph_integrals_noalloc (the integrals with their partial sums in the library's static device scratch)
against osph_integrals (a context buffer allocated on first use), through the C ABI: the same kernels in
the same order, so the results are bit-identical.  Two calls in a row on one GPU, the second on a smaller
state, so that the second reads none of what the first left in the shared array."""
import numpy as np
import pytest

from opensph_amd import gpu, initial

pytestmark = pytest.mark.gpu


def _solver(n, radius, seed):
    st, cfg = initial.fluid_sphere(n_requested=n, radius=radius, sigma_v=3.0, seed=seed)
    st["u"][:] = np.random.default_rng(seed).uniform(0.0, 1.0e4, size=st.n)
    s = gpu.GpuSolver(st, cfg)
    s.step()
    return s


def _same(a, b):
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def test_integrals_noalloc_equal_integrals():
    big = _solver(300_000, 2000.0, 1)    # 1024 blocks of partial sums
    small = _solver(3_000, 500.0, 2)     # 13 blocks
    print(f"[case] integrals noalloc n={big.n} then n={small.n}", flush=True)
    for s in (big, small):
        want = s.integrals()
        got = s.integrals_noalloc()
        _same(got, want)
        assert got["mass"] > 0.0 and np.isfinite(got["energy"])
    # and again on the larger state, after the smaller one used the array
    _same(big.integrals_noalloc(), big.integrals())
    big.close()
    small.close()


def test_integrals_noalloc_of_an_empty_state():
    s = gpu.GpuSolver(initial.ParticleState(0, [initial.basalt()]), initial.RunConfig())
    out = s.integrals_noalloc()
    assert out["mass"] == 0.0 and out["energy"] == 0.0
    s.close()

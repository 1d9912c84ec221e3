"""Reverse-mode gradients of the gas-dynamics PredictorCorrector run (osph_adjoint_forward / _backward)
against central finite differences of the oracle's forward run (the FP64 CPU restatement of the
reference), for the initial velocities, positions, densities and energies and the Tillotson parameters;
independence of the checkpoint interval; mixed precision against FP64.

J = kinetic energy of the outer 20 % of the particles (by final distance from the centre of mass),
the set fixed from the unperturbed run, after STEPS steps of DT."""
import numpy as np
import pytest

from opensph_amd import gpu, initial

pytestmark = pytest.mark.gpu

STEPS = 7
DT = 0.01
EVERY = 3


def _case():
    return initial.gas_collision(n_total=600, radius=1000.0, speed=100.0)


def _outer(st_final, frac=0.8):
    r = st_final["r"][:, :3]
    com = np.sum(st_final["m"][:, None] * r, axis=0) / np.sum(st_final["m"])
    d = np.linalg.norm(r - com, axis=1)
    return d > np.quantile(d, frac)


def _J(st, S):
    return 0.5 * float(np.sum(st["m"][S] * np.sum(st["v"][S, :3] ** 2, axis=1)))


def _gpu_gradients(st, cfg, precision="fp64", every=EVERY, S=None):
    g = gpu.GpuSolver(st, cfg, precision=precision)
    g.adjoint_forward([DT] * STEPS, every)
    fin = g.download(st)
    S = _outer(fin) if S is None else S
    seed_v = np.where(S[:, None], st["m"][:, None] * fin["v"][:, :3], 0.0)
    out = g.adjoint_backward(seed_v=seed_v)
    back = g.download(st)
    g.close()
    return out, S, fin, back


@pytest.fixture(scope="module")
def fp64_run():
    st, cfg = _case()
    print(f"[case] adjoint gas collision n={st.n} steps={STEPS} every={EVERY}", flush=True)
    out, S, fin, back = _gpu_gradients(st, cfg)
    return st, cfg, out, S, fin, back


def _fd(oracle, st, cfg, S, perturb, eps):
    plus, minus = st.copy(), st.copy()
    perturb(plus, +eps)
    perturb(minus, -eps)
    jp = _J(oracle.run_pc_fixed(plus, cfg, [DT] * STEPS), S)
    jm = _J(oracle.run_pc_fixed(minus, cfg, [DT] * STEPS), S)
    return (jp - jm) / (2.0 * eps)


def _components(grad, count, rng):
    flat = np.abs(grad).ravel()
    big = list(np.argsort(flat)[-count:])
    others = rng.choice(np.flatnonzero(flat > 1e-3 * flat.max()), size=count, replace=False)
    return [np.unravel_index(int(k), grad.shape) for k in big + list(others)]


def test_forward_matches_oracle_and_state_returns(fp64_run, oracle):
    st, cfg, out, S, fin, back = fp64_run
    ref = oracle.run_pc_fixed(st, cfg, [DT] * STEPS)
    for k in ("r", "v", "rho", "u"):
        a, b = fin[k], ref[k]
        assert np.linalg.norm(a - b) <= 1e-10 * np.linalg.norm(b), k
    assert out["clamped"] == 0
    for k in ("r", "v", "rho", "u"):
        assert np.array_equal(back[k], st[k]), k  # the resident state is the initial state again


@pytest.mark.parametrize("what,eps", [("v", 1e-4), ("r", 1e-3), ("rho", 1e-3), ("u", 1e-2)])
def test_gradient_matches_central_differences(fp64_run, oracle, what, eps):
    st, cfg, out, S, fin, back = fp64_run
    grad = out[what]
    rng = np.random.default_rng(3)
    comps = _components(grad, 4, rng)
    scale = np.abs(grad).max()
    for c in comps:
        def perturb(s, e, c=c):
            s[what][c] += e
        fd = _fd(oracle, st, cfg, S, perturb, eps)
        assert abs(fd - grad[c]) <= 2e-5 * scale, (what, c, fd, grad[c])


def test_eos_parameter_gradient_matches_central_differences(fp64_run, oracle):
    st, cfg, out, S, fin, back = fp64_run
    for q, name in enumerate(("a", "b", "A")):
        theta = st.materials[0][name]
        rel = 1e-6

        def perturb(s, e, name=name, theta=theta):
            s.materials = [dict(m) for m in s.materials]
            s.materials[0][name] = theta * (1.0 + e)

        fd = _fd(oracle, st, cfg, S, perturb, rel)  # dJ/d(log theta)
        assert abs(fd - out["eos"][0, q] * theta) <= 1e-4 * abs(fd) + 1e-9 * _J(fin, S), (name, fd,
                                                                                          out["eos"][0, q] * theta)


def test_checkpoint_interval_does_not_change_gradients(fp64_run):
    st, cfg, out, S, fin, back = fp64_run
    for every in (1, STEPS):
        other, _, _, _ = _gpu_gradients(st, cfg, every=every, S=S)
        for k in ("r", "v", "rho", "u", "eos"):
            assert np.array_equal(other[k], out[k]), (every, k)


def test_mixed_precision_gradients_close_to_fp64(fp64_run):
    st, cfg, out, S, fin, back = fp64_run
    mixed, _, _, _ = _gpu_gradients(st, cfg, precision="mixed", S=S)
    assert mixed["clamped"] == 0
    for k in ("r", "v", "rho", "u"):
        err = np.linalg.norm(mixed[k] - out[k]) / np.linalg.norm(out[k])
        assert err < 1e-3, (k, err)


def test_mixed_adjoint_matches_central_differences(fp64_run, oracle):
    """The mixed adjoint (FP32 pair terms with FP64 separations and sums, FP64 elsewhere) against the
    oracle's central differences, for the initial velocities and densities."""
    st, cfg, out, S, fin, back = fp64_run
    mixed, _, _, _ = _gpu_gradients(st, cfg, precision="mixed", S=S)
    rng = np.random.default_rng(5)
    for what, eps in (("v", 1e-4), ("rho", 1e-3)):
        grad = mixed[what]
        scale = np.abs(grad).max()
        for c in _components(grad, 3, rng):
            def perturb(s, e, c=c, what=what):
                s[what][c] += e
            fd = _fd(oracle, st, cfg, S, perturb, eps)
            assert abs(fd - grad[c]) <= 5e-5 * scale, (what, c, fd, grad[c])

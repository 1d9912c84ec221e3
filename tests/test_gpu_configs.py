"""The five workload configurations at the checks they state, against the oracle (the FP64 CPU
restatement of the reference) or the quantities the mathematics fixes:

* the 1M-particle impact in mixed precision: accelerations, density and energy rates against the oracle
  at <= 1e-5 (relative L2), on the perturbed exercised impact and on the real impact once the bodies
  touch, with the neighbour counts equal;
* the 8M-particle Barnes-Hut N-body (theta 0.5, quadrupole, leaf 25, M4 softening h = 10 m): both
  precisions against the FP64 direct sum over all sources on a seeded 64k-target sample, the GPU direct
  sum pinned against the oracle's on a sub-sample;
* the impact ensemble: members drawn from the ensemble's laws (log-uniform sizes, angles, speeds, radius
  ratios), stepped as one batch and each checked against the oracle's predictor-corrector run;
* the differentiated gas collision: a 10k-particle instance, 50 steps, a checkpoint every 10, J = kinetic
  energy of the particles beyond 1.2 R; the FP64 adjoint against central differences at eps = 1e-6 and
  the mixed adjoint against the FP64 one.
"""
import numpy as np
import pytest

from conftest import rel_l2
from opensph_amd import gpu, initial

pytestmark = pytest.mark.gpu

ACCEL_TOL = 1e-5  # the mixed path's bar on accelerations, density and energy rates


def _mixed_vs_oracle(oracle, st, cfg):
    g = gpu.GpuSolver(st, cfg, precision="mixed")
    g.integrate()
    a = g.download(st)
    g.close()
    b = oracle.integrate(st, cfg)
    assert np.array_equal(a["ncnt"], b["ncnt"])
    errs = {"dv": rel_l2(a["dv"][:, :3], b["dv"][:, :3]), "drho": rel_l2(a["drho"], b["drho"]),
            "du": rel_l2(a["du"], b["du"])}
    assert np.linalg.norm(b["dv"][:, :3]) > 0.0
    for k, e in errs.items():
        assert e <= ACCEL_TOL, (k, errs)
    return errs


@pytest.mark.parametrize("name", ["impact_exercised", "fluid_exercised"])
def test_mixed_derivatives_match_oracle(oracle, name):
    from cases import exercised_fluid, exercised_impact

    st, cfg = {"impact_exercised": exercised_impact, "fluid_exercised": exercised_fluid}[name]()
    print(f"[case] mixed vs oracle {name} n={st.n}", flush=True)
    _mixed_vs_oracle(oracle, st, cfg)


@pytest.mark.slow
def test_mixed_full_impact_in_contact_matches_oracle(oracle):
    """The 1,062,568-particle impact stepped in mixed precision until two steps past the time the bodies
    touch; at that state the mixed derivatives against the oracle's, and the state against a FP64 run
    over the same steps."""
    st, cfg = initial.impact()
    tc = initial.contact_time(st)
    print(f"[case] mixed full impact n={st.n} contact t={tc:.4f}", flush=True)
    runs = {}
    for prec in ("fp64", "mixed"):
        g = gpu.GpuSolver(st, cfg, precision=prec)
        t, past = 0.0, 0
        while t < tc or past < 2:
            t += g.step()[0]
            past += t >= tc
        runs[prec] = g.download(st)
        g.close()
    errs = _mixed_vs_oracle(oracle, runs["mixed"], cfg)
    print(f"[result] mixed vs oracle in contact: {errs}", flush=True)
    touched = np.count_nonzero(np.any(runs["mixed"]["dv"][:, :3] != 0.0, axis=1))
    assert touched > 100
    assert rel_l2(runs["mixed"]["r"][:, :3], runs["fp64"]["r"][:, :3]) < 1e-9
    assert rel_l2(runs["mixed"]["v"][:, :3], runs["fp64"]["v"][:, :3]) < 1e-7


# ---- the N-body configuration --------------------------------------------------------------------------


def _nbody_check(n, sample_size, oracle_targets, seed=7):
    from opensph_amd.gravity import G_SI, GpuGravity
    from oracle.bridge import GravityOracle

    r, m = initial.nbody_sphere(n, seed=seed)
    sample = np.random.default_rng(3).choice(n, sample_size, replace=False).astype(np.uint32)
    acc = {}
    direct = None
    for prec in ("fp64", "mixed"):
        g = GpuGravity(r, m, solver="barnes_hut", kernel="m4", theta=0.5, order=2, leaf_size=25, precision=prec)
        g.build()
        a, stats = g.eval_self()
        assert stats["nodes_approx"] > 0 and stats["nodes_exact"] > 0
        if direct is None:
            direct = g.direct(sample)
        acc[prec] = a
        g.close()
    # the GPU direct sum against the oracle's on a sub-sample of the targets
    sub = sample[:oracle_targets]
    want = GravityOracle().direct(r, m, G_SI, targets=sub)
    derr = np.linalg.norm(direct[:oracle_targets] - want, axis=1) / np.linalg.norm(want, axis=1)
    assert derr.max() < 1e-11, derr.max()
    rms = np.sqrt(np.mean(np.sum(direct ** 2, axis=1)))
    out = {}
    for prec, a in acc.items():
        diff = np.linalg.norm(a[sample] - direct, axis=1)
        rel = diff / np.linalg.norm(direct, axis=1)
        out[prec] = (float(np.median(rel)), float(diff.max() / rms))
        # quadrupole at theta 0.5: the measured 8M figures are a median of 3.9e-4 and 1.6e-3 of the rms
        assert np.median(rel) < 1e-3 and diff.max() < 5e-3 * rms, (prec, out[prec])
    scale = np.linalg.norm(acc["fp64"], axis=1).max()
    assert np.linalg.norm(acc["mixed"] - acc["fp64"], axis=1).max() < 1e-6 * scale
    return out


def test_nbody_300k_against_oracle_direct_sum():
    print("[case] nbody 300k sphere, 4096 targets", flush=True)
    print(f"[result] {_nbody_check(300_000, 4096, 256)}", flush=True)


@pytest.mark.slow
def test_nbody_8M_against_direct_sum_on_64k_targets():
    print("[case] nbody 8M sphere, 65536 targets", flush=True)
    print(f"[result] {_nbody_check(8_000_000, 65536, 48)}", flush=True)


# ---- the impact ensemble ---------------------------------------------------------------------------------


FIELDS = ("r", "v", "dv", "rho", "drho", "u", "du", "s", "ds", "damage")


def test_ensemble_members_match_oracle(oracle):
    """Four members drawn from the ensemble's laws (sizes log-uniform in [2k, 8k] so that the oracle runs
    in seconds; the bodies placed in contact): three batched steps, FP64 members against the oracle at the
    parity bar and mixed members within the mixed path's bar."""
    ms = initial.impact_ensemble(4, seed=21, n_range=(2e3, 8e3), offset_h=-0.5)
    print("[case] ensemble " + ", ".join(str(st.n) for st, _ in ms), flush=True)
    for prec, tol in (("fp64", 1e-10), ("mixed", ACCEL_TOL)):
        solvers = [gpu.GpuSolver(st, cfg, precision=prec) for st, cfg in ms]
        logs = [[] for _ in ms]
        for _ in range(3):
            dts = [s.dt for s in solvers]
            nxt, crit = gpu.step_batch(solvers)
            for k in range(len(ms)):
                logs[k].append((dts[k], nxt[k], crit[k]))
        for k, ((st, cfg), s) in enumerate(zip(ms, solvers)):
            out = s.download(st)
            ref, rlog = oracle.run_pc(st, cfg, 3)
            for (dt, nx, cr), (odt, onx, ocr) in zip(logs[k], rlog):
                assert nx == pytest.approx(onx, rel=max(tol, 1e-12)), (prec, k)
                if prec == "fp64":
                    assert cr == ocr, (prec, k)
            for f in FIELDS:
                a, b = out[f], ref[f]
                if f == "dv":
                    a, b = a[:, :3], b[:, :3]
                if prec == "mixed" and f == "damage":
                    continue  # a flaw's activation is a threshold decision, which FP32 rounding may move
                if f == "damage":
                    # a flaw activates where strain / eps_min passes 1 (ScalarGradyKippModel::integrate,
                    # Damage.cpp:128-170): a particle at that threshold in the 0.03 s first step and one
                    # that crosses it in the 0.003 s second step end 10x apart.  The device's pow, cbrt
                    # and acos round within an ulp of the C library's, which can move such a particle
                    # across; so can a 1-ulp change of the input h in the oracle itself (member 2 of this
                    # draw: particle 5426, damage 1.7e-3 against 1.5e-4).  At most one particle in a
                    # thousand may be such a case; every other particle is held at the parity bar.
                    off = np.abs(a - b) > tol * max(float(np.abs(b).max()), 1e-300)
                    assert off.sum() <= max(1, st.n // 1000), (prec, k, int(off.sum()))
                    a, b = a[~off], b[~off]
                # mixed: positions, velocities and densities integrate FP32-rounded rates over a few steps
                # on top of FP64 values (the h lane of v is h / 3 div v, with div v's own error); u, S and
                # the rates themselves carry the rates' own error
                bar = 1e-7 if prec == "mixed" and f in ("r", "v", "rho") else tol
                assert rel_l2(a, b) <= bar, (prec, k, f, rel_l2(a, b))
            s.close()


# ---- the differentiated gas collision --------------------------------------------------------------------


ADJ_STEPS, ADJ_EVERY, EPS = 50, 10, 1e-6


@pytest.fixture(scope="module")
def gas10k():
    st, cfg = initial.gas_collision(n_total=10_000, radius=1000.0, speed=100.0)
    g = gpu.GpuSolver(st, cfg)
    g.integrate()
    dt = 0.5 * g.timestep()[0]
    g.close()
    return st, cfg, [dt] * ADJ_STEPS


def _outside(fin, radius=1000.0):
    """The particles beyond 1.2 R of the larger body's centre (the origin) at the end of the run."""
    return np.linalg.norm(fin["r"][:, :3], axis=1) > 1.2 * radius


def _J(st, S):
    return 0.5 * float(np.sum(st["m"][S] * np.sum(st["v"][S, :3] ** 2, axis=1)))


def _forward(st, cfg, dts):
    g = gpu.GpuSolver(st, cfg)
    for dt in dts:
        g.step(dt)
    fin = g.download(st)
    g.close()
    return fin


def _adjoint(st, cfg, dts, precision, S=None):
    g = gpu.GpuSolver(st, cfg, precision=precision)
    g.adjoint_forward(dts, ADJ_EVERY)
    fin = g.download(st)
    S = _outside(fin) if S is None else S
    out = g.adjoint_backward(seed_v=np.where(S[:, None], st["m"][:, None] * fin["v"][:, :3], 0.0))
    g.close()
    return out, S, fin


@pytest.mark.slow
def test_gas_adjoint_10k_against_central_differences(gas10k, oracle):
    st, cfg, dts = gas10k
    print(f"[case] adjoint gas collision n={st.n} steps={ADJ_STEPS} every={ADJ_EVERY}", flush=True)
    out, S, fin = _adjoint(st, cfg, dts, "fp64")
    assert out["clamped"] == 0 and 0 < S.sum() < st.n
    # the recorded FP64 forward run is the oracle's
    ref = oracle.run_pc_fixed(st, cfg, dts)
    for k in ("r", "v", "rho", "u"):
        assert rel_l2(fin[k], ref[k]) < 1e-10, k
    J0 = _J(fin, S)
    rng = np.random.default_rng(8)
    scale = {k: float(np.abs(st[k][:, :3] if k in ("r", "v") else st[k]).max()) for k in ("r", "v", "rho", "u")}

    def central(perturb):
        jp, jm = st.copy(), st.copy()
        perturb(jp, +1.0)
        perturb(jm, -1.0)
        fp, fm = _forward(jp, cfg, dts), _forward(jm, cfg, dts)
        assert np.array_equal(_outside(fp), S) and np.array_equal(_outside(fm), S)
        return (_J(fp, S) - _J(fm, S)) / 2.0

    # directional derivatives along seeded random directions of the whole initial state: the step is
    # eps times each quantity's largest magnitude.  The forward map is only piecewise smooth: the M4
    # table's gradient is linear between its nodes in q^2 (spacing 1e-4) and StandardAV switches on where
    # (v_i - v_j).(r_i - r_j) changes sign, so a central difference across such a kink differs from the
    # one-sided derivative the adjoint takes, by an amount that grows with the number of kinks the step
    # crosses: linearly in eps, not as eps^2.  At eps = 1e-6 this is a few 1e-5 for v, r and u.  A density
    # step moves the pressure through the bulk modulus (A / rho0 ~ 1e7 m^2/s^2), and a relative 1e-6 of it
    # displaces the particles by ~1e-5 h over the run, enough for a large share of the ~350k pair
    # evaluations per step to cross a table node: the difference is 4e-4 at 1e-6, 1.3e-4 at 1e-7 and 9e-6
    # at 1e-8 (measured), converging to the adjoint.  The density is therefore checked at eps = 1e-8 and
    # the convergence from 1e-6 down is asserted; the other quantities at eps = 1e-6.
    steps = {"v": (EPS,), "r": (EPS,), "rho": (1e-6, 1e-7, 1e-8), "u": (EPS,)}
    for what in ("v", "r", "rho", "u"):
        shape = st[what][:, :3].shape if what in ("r", "v") else st[what].shape
        d = rng.standard_normal(shape)
        ad = float(np.sum(out[what] * d))
        errs = []
        for eps in steps[what]:
            step = eps * scale[what]

            def perturb(s, sign, what=what, d=d, step=step):
                if what in ("r", "v"):
                    s[what][:, :3] += sign * step * d
                else:
                    s[what][:] += sign * step * d

            fd = central(perturb) / step
            errs.append(abs(fd - ad))
            print(f"[result] d/d{what}: adjoint {ad:.9e}, central eps {eps:.0e} {fd:.9e} ({(fd - ad) / ad:+.2e})",
                  flush=True)
        assert errs[-1] <= 1e-4 * abs(ad) + 1e-12 * J0 / scale[what], (what, errs, ad)
        assert all(b < a for a, b in zip(errs, errs[1:])), (what, errs)  # the differences converge
    # the Tillotson a, b, A at a relative step eps
    for q, name in enumerate(("a", "b", "A")):
        theta = st.materials[0][name]

        def perturb(s, sign, name=name, theta=theta):
            s.materials = [dict(m) for m in s.materials]
            s.materials[0][name] = theta * (1.0 + sign * EPS)

        fd = central(perturb) / (EPS * theta)
        print(f"[result] d/d{name}: adjoint {out['eos'][0, q]:.9e}, central {fd:.9e}", flush=True)
        assert abs(fd - out["eos"][0, q]) <= 1e-4 * abs(fd) + 1e-12 * J0 / theta, (name, fd, out["eos"][0, q])
    # the mixed-precision adjoint (FP32 pair terms, FP64 sums) against the FP64 adjoint
    mixed, _, _ = _adjoint(st, cfg, dts, "mixed", S=S)
    for k in ("r", "v", "rho", "u"):
        assert rel_l2(mixed[k], out[k]) < 1e-3, k
    assert np.allclose(mixed["eos"], out["eos"], rtol=1e-3, atol=0.0)

"""Throughput of the OpenSPH per-timestep hot path on one MI355X.

Workload: the two-body basalt impact -- target R = 50 km (950,000 requested HCP particles), projectile
R = 5 km (50,000 requested) at 5 km/s and 45 degrees; 1,062,568 particles after lattice generation.
Terms: PressureForce + SolidStressForce + StandardAV + ContinuityEquation + AdaptiveSmoothingLength,
Tillotson EoS, von Mises yielding, scalar Grady-Kipp damage, PredictorCorrector with all four time-step
criteria.  One step = one PredictorCorrector step of every particle with the state resident in HBM.

Prints one JSON line (rank 0).  The CPU baseline is the real OpenSPH core (oracle/_ref/libosph_ref.so,
built from the reference sources) on the host's cores, timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
# Algorithmic FP64 flops per unordered neighbour pair in the pair-sum kernel: the per-term counts of the
# symmetric form (LUT gradient 18, pressure 24, velocity divergence 16, artificial viscosity 60,
# velocity gradient 38, stress divergence 42), without the pair test (done by the list kernel).
FLOP_PER_PAIR = 198.0
# Barnes-Hut evaluation (FP64): a particle pair with the Newtonian branch (difference 3, |r|^2 5, radius
# test 4, 1/sqrt 1, cube 2, mass 1, accumulate 6) and a quadrupole expansion at one particle (difference 3,
# |r|^2 5, 1/sqrt and powers 8, q2 d 15, d q2 d 6, terms 15, monopole 9); counted as in the pair figure.
FLOP_PER_GRAVITY_PAIR = 22.0
FLOP_PER_QUADRUPOLE = 61.0
# HBM bytes per launch of the pair kernel at the default workload and tuning, after contact, from two
# rocprofv3 --pmc passes over this bench (FETCH_SIZE scaled by 2.000, WRITE_SIZE by 1.000: the factors the
# 1 GiB device copy of --calibrate reads).  None where not measured.
MEASURED_TRAFFIC_BYTES = {"pairs_mixed_finalize_correct": (2 * 420926 + 312715) * 1024.0,
                          "pairs_fp64_finalize_correct": (2 * 557645 + 315804) * 1024.0}
# Algorithmic HBM bytes of one pair-kernel launch per particle: each particle's own neighbour record
# (64 B mixed, 128 B FP64), its order, count and group entries and the 92 B it writes (acceleration,
# du, divergence, velocity gradient: 11 doubles, and the neighbour count); plus 4 B per listed neighbour.
# The fused kernels (the PC step's default) add the finalize + corrector + criteria pass, which takes the
# pair sums in registers: 32 doubles read (state, predictions, damage and EoS inputs), 23 written, and
# the material id, 444 B.
PAIR_BYTES_PER_PARTICLE = {"pairs_mixed": 168.0, "pairs_fp64": 232.0,
                           "pairs_mixed_finalize_correct": 168.0 + 444.0,
                           "pairs_fp64_finalize_correct": 232.0 + 444.0}
PAIR_BYTES_PER_NEIGHBOR = 4.0
# VALU wave-instructions per dispatch of the (fused) pair kernels at the 1M impact after contact
# (rocprofv3 --pmc SQ_INSTS_VALU, averaged over 40 dispatches; the finalize + corrector + criteria pass
# included), per 64 pair evaluations (the gather form evaluates each unordered pair from both sides):
# the issue-rate roof counts instructions, not flops.  A CU issues one
# wave64 VALU instruction per cycle (four SIMDs, four cycles each), so the chip's roof is
# 256 CUs x 2.4 GHz = 6.14e11 wave-instructions/s (x 64 lanes x 2 = the 78.6 TFLOP/s FMA peak).
VALU_INSTR_PER_64_PAIR_EVALS = {"pairs_mixed_finalize_correct": 1.833e8 / (2 * 35218390 / 64),
                                "pairs_fp64_finalize_correct": 1.970e8 / (2 * 35218390 / 64)}
VALU_ISSUE_PEAK = 256 * 2.4e9
# The same for the Barnes-Hut evaluation of the 8,000,000-particle N-body sphere (FETCH_SIZE x 2 +
# WRITE_SIZE, KiB per dispatch, averaged over the 2 dispatches with the state in Morton order).
NBODY_TRAFFIC_BYTES = {"eval_mixed": (2 * 2045976 + 245187) * 1024.0, "eval": (2 * 1957327 + 248606) * 1024.0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(st, cfg, dt0, window, max_seconds=25.0):
    """Times the reference core (ThreadPool, all host cores) on PredictorCorrector steps of the same window
    of the workload: it starts from the FP64 GPU run's state at the start of its timed window (in contact)
    with that run's time step, and runs until about max_seconds of CPU wall time; falls back to the oracle
    port on one core."""
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    where = (f"from t = {window['t_start']:.4f} s (after {window['steps_before']} steps from t = 0; the bodies "
             f"touch at t = {window['t_contact']:.4f} s)")
    try:
        from oracle.bridge import RefRun

        if RefRun.available():
            ref = RefRun(st, cfg, nthreads=threads, dt0=dt0)
            ref.step()  # first step (tree and buffers allocated) not timed
            steps, t0 = 0, time.perf_counter()
            while True:
                ref.step()
                steps += 1
                el = time.perf_counter() - t0
                if el > max_seconds or steps >= 10:
                    break
            ref.close()
            return {"value": st.n * steps / el, "unit": "particle-steps/s", "cores": threads, "kind": "reference",
                    "sample": f"{steps} PredictorCorrector steps (after one untimed) of the same {st.n}-particle "
                              f"impact {where}, OpenSPH core built from source (-O2, strict FP), "
                              f"ThreadPool x{threads}"}
        from oracle.bridge import Oracle

        o = Oracle()
        t0 = time.perf_counter()
        o.run_pc(st, cfg, 1, dt0=dt0)
        el = time.perf_counter() - t0
        return {"value": st.n / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                "sample": f"1 PredictorCorrector step of the same {st.n}-particle impact {where}, oracle C port, "
                          f"1 core"}
    except Exception as e:  # the baseline is reported, never required
        log("cpu baseline failed:", e)
        return None


def total_energy(st):
    v2 = np.sum(st["v"][:, :3] ** 2, axis=1)
    return float(np.sum(st["m"] * (0.5 * v2 + st["u"])))


def run_precision(gpu, st, cfg, prec, warmup, steps, t_contact, total_steps=100, tune=None, tuner=False):
    """One precision of the impact from t = 0: untimed PC steps into contact (at least two steps past the
    contact time), the warmup, the timed window, a per-kernel profile over the next steps, then untimed
    steps up to total_steps for the energy drift of the whole run."""
    solver = gpu.GpuSolver(st, cfg, precision=prec)
    if tune:
        solver.tune(**tune)
    t, done, past = 0.0, 0, 0
    while t < t_contact or past < 2:
        dt, _, _ = solver.step()
        t += dt
        done += 1
        past += t >= t_contact
    prelude = done
    tuned = None
    if tuner:
        # launch shapes picked by the autotuner on this state (it leaves the state as it found it)
        from opensph_amd import autotune
        tuned = autotune.autotune(solver, steps=10, log=log)
        log(f"{prec}: autotuned {tuned['settings']} ({tuned['ms_per_step']:.3f} ms/step)")
    for _ in range(warmup):
        dt, _, _ = solver.step()
        t += dt
    done += warmup
    solver.sync()
    start = solver.download(st)
    t_start, dt_start = t, solver.dt
    # the timed steps carry HIP events around the pair-sum launches only (for the roofline); events
    # around every launch would add a few microseconds between kernels
    solver.profile(2)
    dts = []
    t0 = time.perf_counter()
    for _ in range(steps):
        dts.append(solver.step()[0])
    solver.sync()
    elapsed = time.perf_counter() - t0
    prof_pairs = solver.profile_read()
    done += steps
    t += sum(dts)
    end = solver.download(st)
    t_end = t
    # per-kernel breakdown: every launch timed over the next prof_steps steps (one window)
    prof_steps = min(max(steps, 1), 10)
    solver.profile(1)
    for _ in range(prof_steps):
        t += solver.step()[0]
    prof = solver.profile_read()
    solver.profile(False)
    done += prof_steps
    while done < total_steps:
        t += solver.step()[0]
        done += 1
    final = solver.download(st)
    solver.close()
    return {"value": st.n * steps / elapsed, "elapsed": elapsed, "ms_per_step": 1e3 * elapsed / steps,
            "prof": prof, "prof_steps": prof_steps, "prof_pairs": prof_pairs, "start": start, "end": end,
            "final": final, "dt_start": dt_start,
            "window": {"t_contact": t_contact, "steps_before": prelude + warmup, "prelude_steps": prelude,
                       "t_start": t_start, "t_end": t_end, "run_steps": done, "t_run_end": t},
            "tuned": tuned["settings"] if tuned else None}


def compare_precisions(gpu, st, cfg, runs, max_accel_err=1e-5):
    """Acceleration and density-rate error of the mixed pair sums against FP64 on the same state (the FP64
    run at the end of its timed window, in contact), and the energy drift of both runs over the whole run
    from t = 0.  Raises when the state has no pair sums (the bodies are not in contact) or when the mixed
    error exceeds max_accel_err: the headline is only valid at that accuracy."""
    ref_state = runs["fp64"]["end"]
    out = {}
    for prec in ("fp64", "mixed"):
        g = gpu.GpuSolver(ref_state, cfg, precision=prec)
        g.integrate()
        out[prec] = g.download(ref_state)
        g.close()
    a64, amx = out["fp64"]["dv"][:, :3], out["mixed"]["dv"][:, :3]
    n64 = float(np.linalg.norm(a64))
    if not (n64 > 0.0) or not np.isfinite(n64):
        raise RuntimeError("the FP64 accelerations of the window's state are zero or not finite: the bodies are "
                           "not in contact, so the window does not exercise the pair sums")
    e0 = total_energy(st)
    e64, emx = total_energy(runs["fp64"]["final"]), total_energy(runs["mixed"]["final"])
    acc = {"state": "FP64 run at t_end of its timed window",
           "accel_rel_l2": float(np.linalg.norm(amx - a64) / n64),
           "accel_err_over_max": float(np.max(np.linalg.norm(amx - a64, axis=1)) /
                                       np.max(np.linalg.norm(a64, axis=1))),
           "drho_rel_l2": float(np.linalg.norm(out["mixed"]["drho"] - out["fp64"]["drho"]) /
                                np.linalg.norm(out["fp64"]["drho"])),
           "du_rel_l2": float(np.linalg.norm(out["mixed"]["du"] - out["fp64"]["du"]) /
                              max(np.linalg.norm(out["fp64"]["du"]), 1e-300)),
           "particles_accelerated": int(np.count_nonzero(np.any(a64 != 0.0, axis=1))),
           "neighbor_counts_equal": bool(np.array_equal(out["mixed"]["ncnt"], out["fp64"]["ncnt"])),
           "energy_steps": {p: runs[p]["window"]["run_steps"] for p in ("fp64", "mixed")},
           "energy_drift_fp64": (e64 - e0) / e0, "energy_drift_mixed": (emx - e0) / e0,
           "energy_mixed_vs_fp64": (emx - e64) / e0,
           "gate": {"accel_rel_l2_max": max_accel_err, "energy_drift_max": 1e-6}}
    if not acc["accel_rel_l2"] <= max_accel_err:
        raise RuntimeError(f"mixed-precision acceleration error {acc['accel_rel_l2']:.3g} exceeds {max_accel_err:g}")
    if not (abs(acc["energy_drift_mixed"]) <= 1e-6 and abs(acc["energy_drift_fp64"]) <= 1e-6):
        raise RuntimeError(f"energy drift over {acc['energy_steps']} steps exceeds 1e-6: fp64 "
                           f"{acc['energy_drift_fp64']:.3g}, mixed {acc['energy_drift_mixed']:.3g}")
    return acc


def run_nbody(args):
    """Self-gravity N-body: Barnes-Hut (theta 0.5, quadrupole, leaf 25, M4-softened) + LeapFrog, dt = 1 s.
    --precision both (the default) runs FP64 and then the FP32 evaluation from the same start and reports
    the mixed run, with the FP64 run and the acceleration difference beside it."""
    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    runs = {p: nbody_once(args, p) for p in precs}
    out = runs[precs[-1]]
    if len(runs) == 2:
        a64, amx = runs["fp64"].pop("_dv0"), runs["mixed"].pop("_dv0")
        scale = np.linalg.norm(a64, axis=1).max()
        out["fp64_value"] = runs["fp64"]["value"]
        out["fp64_kernel_ms_per_step"] = runs["fp64"]["kernel_ms_per_step"]
        out["fp64_roofline"] = runs["fp64"]["roofline"]
        out["fp64_accuracy"] = runs["fp64"]["accuracy"]
        out["mixed_over_fp64"] = out["value"] / runs["fp64"]["value"]
        out["mixed_vs_fp64_first_eval"] = {
            "err_over_max_max": float(np.linalg.norm(amx - a64, axis=1).max() / scale),
            "rel_l2": float(np.linalg.norm(amx - a64) / np.linalg.norm(a64))}
    out.pop("_dv0", None)
    print(json.dumps(out), flush=True)


def nbody_once(args, prec):
    from opensph_amd import gpu, initial
    from opensph_amd.gravity import GpuGravity

    n = args.nbody_n
    r, m = initial.nbody_sphere(n)
    g = GpuGravity(r, m, solver="barnes_hut", kernel="m4", theta=0.5, order=2, leaf_size=25, precision=prec)
    # the accelerations of the first evaluation, for the FP32 / FP64 comparison (untimed)
    g.build()
    dv0, _ = g.eval_self()
    g.set_particles(r, m)
    g.leapfrog(1.0, args.warmup)
    g.sync()
    g.profile(True)
    t0 = time.perf_counter()
    g.leapfrog(1.0, args.steps)
    g.sync()
    elapsed = time.perf_counter() - t0
    prof = g.profile_read()
    g.profile(False)
    # One evaluation at the final positions, outside the timed region: its interaction counts give the
    # eval kernel's roofline, and its accelerations are checked against the direct sum on a 64k-target
    # sample.  (The stored dv is half a drift older than the positions after the last step.)
    g.build()
    g.profile(True)
    dv_now, st = g.eval_self()
    kname = "eval_mixed" if prec == "mixed" else "eval"
    eval_ms = g.profile_read().get(kname, (float("nan"), 1))[0]
    g.profile(False)
    sample = np.random.default_rng(3).choice(n, min(n, 65536), replace=False).astype(np.uint32)
    bh = dv_now[sample]
    direct = g.direct(sample)
    rel = np.linalg.norm(bh - direct, axis=1) / np.linalg.norm(direct, axis=1)
    rms = np.sqrt(np.mean(np.sum(direct ** 2, axis=1)))
    g.close()
    name, cus = gpu.device_info()
    peak = FP32_PEAK_TFLOPS if prec == "mixed" else FP64_PEAK_TFLOPS
    out = {"metric": "particle-steps/s, 8M N-body Barnes-Hut + LeapFrog", "value": n * args.steps / elapsed,
           "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
           "dtype": "f32+f64" if prec == "mixed" else "f64", "data": "synthetic",
           "config": {"workload": "uniform sphere R 100 km, equal masses, M4 softening h 10 m, theta 0.5, "
                                  "quadrupole, leaf 25, LeapFrog dt 1 s", "particles": n, "precision": prec,
                      "device": name},
           "kernel_ms_per_step": {k: v[0] / max(args.steps, 1) for k, v in sorted(prof.items())},
           "accuracy": {"sample": int(sample.size), "rel_err_median": float(np.median(rel)),
                        "rel_err_max": float(rel.max()),
                        "err_over_rms_max": float(np.max(np.linalg.norm(bh - direct, axis=1)) / rms)},
           "interactions": {"nodes_approx": int(st["nodes_approx"]), "leaves_exact": int(st["nodes_exact"]),
                            "expansions": int(st["expansions"]), "pairs": int(st["pairs"]), "eval_ms": eval_ms},
           "roofline": {"bound": "valu", "kernel": kname,
                        "achieved": (FLOP_PER_GRAVITY_PAIR * st["pairs"] + FLOP_PER_QUADRUPOLE * st["expansions"])
                        / (eval_ms * 1e-3) / 1e12,
                        "peak": peak, "unit": "TFLOP/s",
                        "traffic": NBODY_TRAFFIC_BYTES.get(kname) if n == 8_000_000 else None},
           "_dv0": dv0}
    out["roofline"]["frac"] = out["roofline"]["achieved"] / peak
    return out


def run_ensemble(args):
    """Impact ensemble: all members stepped together (osph_step_pc_batch) against one after another."""
    from opensph_amd import gpu, initial

    # set-up on the device: lattices, placement and the Grady-Kipp flaws (osph_setup_spheres, osph_set_flaws)
    t0 = time.perf_counter()
    specs = initial.ensemble_specs(args.ensemble_n)
    solvers = [initial.device_impact(precision="mixed", flaw_seed=1000 * k, **spec) for k, spec in enumerate(specs)]
    for s in solvers:
        s.sync()
    t_setup = time.perf_counter() - t0
    total = sum(s.n for s in solvers)
    log(f"ensemble of {len(solvers)} members, {total} particles, set up on the device in {t_setup:.1f}s")
    for _ in range(args.warmup):
        gpu.step_batch(solvers)
    for s in solvers:
        s.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gpu.step_batch(solvers)
    for s in solvers:
        s.sync()
    batched = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for s in solvers:
            s.step()
    for s in solvers:
        s.sync()
    sequential = time.perf_counter() - t0
    name, _ = gpu.device_info()
    out = {"metric": "particle-steps/s, impact ensemble (mixed precision, per-member time steps, one launch per "
                     "stage for all members)",
           "value": total * args.steps / batched, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * batched / args.steps, "higher_is_better": True,
           "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": "impact ensemble", "members": len(solvers), "particles": total, "device": name},
           "sequential_value": total * args.steps / sequential, "setup_s": t_setup}
    print(json.dumps(out), flush=True)


def run_adjoint(args):
    """Gradient of the outer kinetic energy of a two-sphere gas collision (PressureForce + StandardAV +
    ContinuityEquation, Tillotson) with respect to the initial state and the Tillotson a, b, A: forward
    run with checkpoints, then the checkpointed reverse sweep (one replay plus one reverse step per step)."""
    from opensph_amd import gpu, initial

    t0 = time.perf_counter()
    st, cfg = initial.gas_collision(n_total=args.adjoint_n, radius=1000.0, speed=100.0)
    log(f"gas collision of {st.n} particles set up in {time.perf_counter() - t0:.1f}s")
    steps, every = args.adjoint_steps, args.adjoint_every
    out = {}
    for prec in ("fp64", "mixed"):
        g = gpu.GpuSolver(st, cfg, precision=prec)
        g.integrate()  # the sound speed the Courant criterion needs
        dt = 0.5 * g.timestep()[0]
        g.upload(st)  # back to the initial state (zero derivatives)
        dts = [dt] * steps
        g.adjoint_forward(dts[:2], 1)  # warm-up (allocations, kernel loads)
        g.adjoint_backward(seed_v=np.zeros((st.n, 3)))
        g.sync()
        t0 = time.perf_counter()
        g.adjoint_forward(dts, every)
        g.sync()
        fwd = time.perf_counter() - t0
        fin = g.download(st)
        r = fin["r"][:, :3]
        R = float(np.max(np.linalg.norm(st["r"][:, :3] - st["r"][:, :3].mean(axis=0), axis=1)))
        com = np.sum(fin["m"][:, None] * r, axis=0) / np.sum(fin["m"])
        S = np.linalg.norm(r - com, axis=1) > 0.5 * R
        seed_v = np.where(S[:, None], st["m"][:, None] * fin["v"][:, :3], 0.0)
        t0 = time.perf_counter()
        grads = g.adjoint_backward(seed_v=seed_v)
        g.sync()
        bwd = time.perf_counter() - t0
        g.close()
        out[prec] = {"forward_s": fwd, "backward_s": bwd, "grads": grads, "dt": dt, "outer": int(S.sum())}
    a, b = out["mixed"]["grads"], out["fp64"]["grads"]
    rel = {k: float(np.linalg.norm(a[k] - b[k]) / np.linalg.norm(b[k])) for k in ("v", "r", "rho", "u")}
    name, _ = gpu.device_info()
    main_run = out["mixed"]
    total = main_run["forward_s"] + main_run["backward_s"]
    res = {"metric": "particle-steps/s, gas-collision gradient (forward + checkpointed reverse sweep)",
           "value": st.n * steps / total, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps, "warmup": 2,
           "ms_per_step": 1e3 * total / steps, "higher_is_better": True, "dtype": "f32+f64",
           "data": "synthetic",
           "config": {"workload": "two basalt spheres R 1 km and 0.5 km, 100 m/s, PressureForce + StandardAV + "
                                  "ContinuityEquation, Tillotson, constant h, dt fixed at half the Courant step",
                      "particles": st.n, "checkpoint_every": every, "device": name},
           "objective": f"kinetic energy of the {main_run['outer']} particles beyond R/2 of the centre of mass",
           "forward_s": {p: out[p]["forward_s"] for p in out}, "backward_s": {p: out[p]["backward_s"] for p in out},
           "fp64_value": st.n * steps / (out["fp64"]["forward_s"] + out["fp64"]["backward_s"]),
           "mixed_vs_fp64_gradient_rel_l2": rel, "clamped": {p: out[p]["grads"]["clamped"] for p in out},
           "dJ_dtheta_mixed": out["mixed"]["grads"]["eos"][0].tolist(),
           "dJ_dtheta_fp64": out["fp64"]["grads"]["eos"][0].tolist()}
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=80)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"],
                    help="both: time the FP64 drop-in and the mixed-precision form; value = mixed")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--total-steps", type=int, default=100,
                    help="impact: steps from t = 0 over which the energy drift is reported (untimed steps after "
                         "the timed window and its profile fill up to this count)")
    ap.add_argument("--tune", action="append", default=[],
                    help="launch-shape setting key=value passed to osph_tune (e.g. pairs_lds=40960)")
    ap.add_argument("--gravity", type=int, default=0, choices=[0, 1, 2],
                    help="impact with self-gravity (GravitySolver): 1 direct sum, 2 Barnes-Hut theta 0.5 octupole")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the launch shapes with opensph_amd.autotune before the warmup")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run the 1 GiB device copy (HBM roof, FETCH_SIZE/WRITE_SIZE calibration)")
    ap.add_argument("--workload", default="impact", choices=["impact", "nbody", "ensemble", "adjoint"],
                    help="impact: the SPH headline; nbody: self-gravity Barnes-Hut + LeapFrog; "
                         "ensemble: batched impact ensemble; adjoint: gas-collision gradient")
    ap.add_argument("--adjoint-n", type=int, default=500_000)
    ap.add_argument("--adjoint-steps", type=int, default=50)
    ap.add_argument("--adjoint-every", type=int, default=10)
    ap.add_argument("--nbody-n", type=int, default=8_000_000)
    ap.add_argument("--ensemble-n", type=int, default=256)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs the hot path on a single MI355X; "
              "--gpus must be 1", file=sys.stderr)
        sys.exit(2)

    from opensph_amd import gpu, initial

    if args.workload == "nbody":
        run_nbody(args)
        return
    if args.workload == "adjoint":
        return run_adjoint(args)
    if args.workload == "ensemble":
        run_ensemble(args)
        return

    if args.calibrate:
        nbytes, iters = 1 << 30, 10
        gbs = gpu.stream_copy(nbytes, iters)
        print(json.dumps({"copy_bytes": nbytes, "iters": iters + 1, "read_plus_write_GBs": gbs}), flush=True)
        return

    t0 = time.perf_counter()
    st, cfg = initial.impact()
    if args.gravity:
        import dataclasses
        cfg = dataclasses.replace(cfg, gravity=args.gravity)
    t_contact = initial.contact_time(st)
    log(f"setup {st.n} particles in {time.perf_counter() - t0:.1f}s; the bodies touch at t = {t_contact:.4f} s")
    name, cus = gpu.device_info()
    precisions = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    runs = {}
    for prec in precisions:
        tune = dict((kv.split("=")[0], int(kv.split("=")[1])) for kv in args.tune)
        runs[prec] = run_precision(gpu, st, cfg, prec, args.warmup, args.steps, t_contact, args.total_steps, tune,
                                   args.autotune)
        log(f"{prec}: {runs[prec]['value']:.4g} particle-steps/s, {runs[prec]['ms_per_step']:.3f} ms/step")
    main_prec = precisions[-1]
    r = runs[main_prec]
    value = r["value"]
    elapsed = r["elapsed"]

    # unordered pairs in the timed window: the neighbour counts of its first and last evaluation
    nb_start = float(np.mean(r["start"]["ncnt"]))
    nb_end = float(np.mean(r["end"]["ncnt"]))
    nb_mean = 0.5 * (nb_start + nb_end)
    pairs = nb_mean * st.n / 2.0
    kernel = "pairs_mixed" if main_prec == "mixed" else "pairs_fp64"
    prof_pairs = r["prof_pairs"]
    if kernel not in prof_pairs and kernel + "_finalize_correct" in prof_pairs:
        # the PC step fuses the finalize, corrector and criteria into the pair kernel: the roofline then
        # holds the pair flops against the fused kernel's time (a lower bound of the pair part's rate)
        kernel += "_finalize_correct"
    ms, launches = prof_pairs.get(kernel, (float("nan"), 1))
    avg_ms = ms / max(launches, 1)
    achieved = FLOP_PER_PAIR * pairs / (avg_ms * 1e-3) / 1e12
    peak = FP32_PEAK_TFLOPS if main_prec == "mixed" else FP64_PEAK_TFLOPS
    alg_bytes = st.n * (PAIR_BYTES_PER_PARTICLE[kernel] + PAIR_BYTES_PER_NEIGHBOR * nb_mean)
    traffic = MEASURED_TRAFFIC_BYTES.get(kernel) if st.n == 1062568 else None

    def per_step(run):
        # every launch of the profiled window, divided by its step count: true per-step times
        return {k: v[0] / run["prof_steps"] for k, v in sorted(run["prof"].items())}

    steps_ms = per_step(r)
    roofline = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": traffic, "kernel": kernel,
                "kernel_ms_per_launch": avg_ms, "launches_timed": launches,
                "flop_per_pair": FLOP_PER_PAIR, "pairs_per_launch": pairs,
                "hbm": {"algorithmic_bytes": alg_bytes, "achieved_GBs": alg_bytes / (avg_ms * 1e-3) / 1e9,
                        "peak_GBs": HBM_PEAK_GBS, "frac": alg_bytes / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
                "dominant_kernel": max(steps_ms.items(), key=lambda kv: kv[1])[0] if steps_ms else kernel,
                "kernel_ms_per_step": steps_ms,
                "kernel_launches_per_step": {k: v[1] / r["prof_steps"] for k, v in sorted(r["prof"].items())},
                "profiled_steps": r["prof_steps"]}
    if main_prec == "mixed":
        # FP32 without packed instructions (disabled in this build) issues at the FP64 rate
        roofline["frac_of_unpacked_fp32"] = achieved / FP64_PEAK_TFLOPS
    if kernel in VALU_INSTR_PER_64_PAIR_EVALS:
        # the instructions the kernel issues (measured count per pair evaluation x this launch's pairs)
        # over the launch time, against the chip's VALU issue rate
        instr = VALU_INSTR_PER_64_PAIR_EVALS[kernel] * 2.0 * pairs / 64.0
        roofline["valu_issue"] = {"instr_per_64_pair_evals": VALU_INSTR_PER_64_PAIR_EVALS[kernel],
                                  "instr_per_launch": instr, "peak_instr_per_s": VALU_ISSUE_PEAK,
                                  "frac": instr / (avg_ms * 1e-3) / VALU_ISSUE_PEAK}
    accuracy = None
    if "fp64" in runs and "mixed" in runs:
        accuracy = compare_precisions(gpu, st, cfg, runs)

    # the reference driver's self-gravity is the direct sum: O(N^2) per step at 1M particles, not a
    # bounded sample, so the gravity variants report no CPU baseline
    base_run = runs["fp64"] if "fp64" in runs else r
    baseline = None if (args.no_cpu or args.gravity) else cpu_baseline(base_run["start"], cfg, base_run["dt_start"],
                                                                        base_run["window"])
    out = {
        "metric": "particle-steps/s, 1M impact: mixed vs FP64 GPU at <=1e-5 accel err",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_prec == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "basalt two-body impact (R 50 km + 5 km, 5 km/s, 45 deg), solid stress + AV + "
                               "continuity + adaptive h, von Mises + Grady-Kipp damage, predictor-corrector",
                   "particles": st.n, "precision": main_prec, "device": name, "compute_units": cus,
                   "self_gravity": {0: None, 1: "direct sum", 2: "Barnes-Hut theta 0.5 octupole leaf 25"}[args.gravity],
                   "neighbors_mean": nb_mean},
        "window": r["window"],
        "fp64_value": runs["fp64"]["value"] if "fp64" in runs else None,
        "fp64_ms_per_step": runs["fp64"]["ms_per_step"] if "fp64" in runs else None,
        "fp64_window": runs["fp64"]["window"] if "fp64" in runs else None,
        "fp64_kernel_ms_per_step": per_step(runs["fp64"]) if "fp64" in runs else None,
        "mixed_over_fp64": (runs["mixed"]["value"] / runs["fp64"]["value"]) if len(runs) == 2 else None,
        "accuracy": accuracy,
        "autotuned": {p: runs[p]["tuned"] for p in runs} if args.autotune else None,
        "roofline": roofline,
        "cpu_baseline": baseline,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

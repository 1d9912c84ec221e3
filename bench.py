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
# Algorithmic FP64 flops per unordered neighbour pair in the pair-sum kernel: the per-term counts of the
# symmetric form (LUT gradient 18, pressure 24, velocity divergence 16, artificial viscosity 60,
# velocity gradient 38, stress divergence 42), without the pair test (done by the list kernel).
FLOP_PER_PAIR = 198.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(st, cfg, max_seconds=25.0):
    """Times the reference core (ThreadPool, all host cores) on PredictorCorrector steps of the same
    workload until about max_seconds of CPU wall time; falls back to the oracle port on one core."""
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    try:
        from oracle.bridge import RefRun

        if RefRun.available():
            ref = RefRun(st, cfg, nthreads=threads)
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
                              f"impact, OpenSPH core built from source (-O2, strict FP), ThreadPool x{threads}"}
        from oracle.bridge import Oracle

        o = Oracle()
        sub = st
        t0 = time.perf_counter()
        o.run_pc(sub, cfg, 1)
        el = time.perf_counter() - t0
        return {"value": sub.n / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                "sample": f"1 PredictorCorrector step of the same {sub.n}-particle impact, oracle C port, 1 core"}
    except Exception as e:  # the baseline is reported, never required
        log("cpu baseline failed:", e)
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=80)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run the 1 GiB device copy (HBM roof, FETCH_SIZE/WRITE_SIZE calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs the hot path on a single MI355X; "
              "--gpus must be 1", file=sys.stderr)
        sys.exit(2)

    from opensph_amd import gpu, initial

    if args.calibrate:
        nbytes, iters = 1 << 30, 10
        gbs = gpu.stream_copy(nbytes, iters)
        print(json.dumps({"copy_bytes": nbytes, "iters": iters + 1, "read_plus_write_GBs": gbs}), flush=True)
        return

    t0 = time.perf_counter()
    st, cfg = initial.impact()
    log(f"setup {st.n} particles in {time.perf_counter() - t0:.1f}s")
    solver = gpu.GpuSolver(st, cfg, precision=args.precision)
    name, cus = gpu.device_info()
    for _ in range(args.warmup):
        solver.step()
    solver.sync()
    solver.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.step()
    solver.sync()
    elapsed = time.perf_counter() - t0
    prof = solver.profile_read()
    solver.profile(False)
    stats = solver.integrate()  # neighbour statistics of the final state (outside the timed region)
    value = st.n * args.steps / elapsed

    pairs = stats["neighbors_mean"] * st.n / 2.0
    kernel = "pairs_fp64"
    ms, launches = prof.get(kernel, (float("nan"), 1))
    avg_ms = ms / max(launches, 1)
    achieved = FLOP_PER_PAIR * pairs / (avg_ms * 1e-3) / 1e12
    dominant = max(prof.items(), key=lambda kv: kv[1][0])[0] if prof else kernel
    roofline = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": kernel,
                "kernel_ms": avg_ms, "dominant_kernel": dominant,
                "kernel_ms_per_step": {k: v[0] / max(v[1], 1) for k, v in sorted(prof.items())}}

    baseline = None if args.no_cpu else cpu_baseline(st, cfg)
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
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "basalt two-body impact (R 50 km + 5 km, 5 km/s, 45 deg), solid stress + AV + "
                               "continuity + adaptive h, von Mises + Grady-Kipp damage, predictor-corrector",
                   "particles": st.n, "precision": args.precision, "device": name, "compute_units": cus,
                   "neighbors_mean": stats["neighbors_mean"]},
        "roofline": roofline,
        "cpu_baseline": baseline,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

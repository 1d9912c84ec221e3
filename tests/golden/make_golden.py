"""Writes the golden vectors under tests/golden/ from the real OpenSPH core.

Runs only where the reference core has been built (oracle/_ref/libosph_ref.so, built by
`make -C oracle ref` from the reference sources).  Each fixture holds the inputs and the reference's
outputs (SymmetricSolver<3>::integrate, then PredictorCorrector steps with its MultiCriterion):

  fluid.npz        ~2.1k-particle basalt fluid sphere, perturbed densities/energies/velocities
  impact.npz       ~1.8k-particle two-body solid impact exercising every material branch
  fluid_pairs.npz  sha256 of the sorted KdTree pair set of the 10,648-particle sphere
"""
import dataclasses
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from cases import exercised_fluid, exercised_impact  # noqa: E402
from oracle.bridge import RefRun  # noqa: E402
from opensph_amd import initial  # noqa: E402

IN_FIELDS = ["r", "v", "rho", "u", "s", "damage", "m", "reduce", "flag", "mat", "eps_min", "m_zero", "growth",
             "n_flaws"]
INTEG_FIELDS = ["dv", "du", "divv", "drho", "p", "cs", "ds", "gradv", "ddamage", "reduce", "s", "ncnt"]
STEP_FIELDS = ["r", "v", "rho", "u", "s", "damage", "dv", "drho", "du", "ds"]


def write_case(name, st, cfg, steps):
    ref = RefRun(st, cfg)
    integ = ref.integrate()
    ref.close()
    ref = RefRun(st, cfg)
    dts = [ref.step() for _ in range(steps)]
    fin = ref.read()
    ref.close()
    out = {"n": np.int64(st.n), "dts": np.array(dts, dtype=np.float64)}
    for k in IN_FIELDS:
        out["in_" + k] = st[k]
    for k in INTEG_FIELDS:
        out["integ_" + k] = integ[k]
    for k in STEP_FIELDS:
        out["steps_" + k] = fin[k]
    meta = {"cfg": dataclasses.asdict(cfg), "materials": st.materials, "steps": steps}
    out["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, name), **out)
    print(name, st.n, os.path.getsize(os.path.join(HERE, name)))


def main():
    st, cfg = exercised_fluid(seed=11)
    st = st.subset(np.arange(min(st.n, 2200)))
    write_case("fluid.npz", st, cfg, 4)
    st, cfg = exercised_impact(n_target=1500, n_projectile=150, seed=21)
    write_case("impact.npz", st, cfg, 4)
    st, cfg = initial.fluid_sphere()
    ref = RefRun(st, cfg)
    pi, pj = ref.pairs()
    keys = (pi.astype(np.uint64) << np.uint64(32)) | pj.astype(np.uint64)
    np.savez_compressed(os.path.join(HERE, "fluid_pairs.npz"), count=np.int64(len(keys)),
                        sha256=np.array(hashlib.sha256(keys.tobytes()).hexdigest()))
    print("fluid_pairs.npz", len(keys))


if __name__ == "__main__":
    main()

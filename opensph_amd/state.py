"""Host-side particle state and run configuration for the OpenSPH hot path.

`ParticleState` mirrors the quantities the reference keeps in its `Storage` for the SymmetricSolver /
PredictorCorrector path (core/quantities/Storage.h:242-575, quantity ids core/quantities/QuantityIds.h),
as named numpy arrays in storage order.  Vector quantities keep the reference's 4-lane layout
(x, y, z, h) (core/objects/geometry/Vector.h:21-26); traceless tensors keep the five independent
components (xx, yy, xy, xz, yz) (core/objects/geometry/TracelessTensor.h:36-45); symmetric tensors are
(xx, yy, zz, xy, xz, yz).
"""
from __future__ import annotations

import dataclasses
import sys
from typing import Dict, List

import numpy as np

INFTY = sys.float_info.max  # core/math/MathUtils.h:38 (std::numeric_limits<double>::max())

# name -> (dtype, components)
FIELDS: Dict[str, tuple] = {
    "r": (np.float64, 4),        # POSITION value: x, y, z, h
    "v": (np.float64, 4),        # POSITION dt
    "dv": (np.float64, 4),       # POSITION d2t
    "rho": (np.float64, 1), "drho": (np.float64, 1),
    "u": (np.float64, 1), "du": (np.float64, 1),
    "s": (np.float64, 5), "ds": (np.float64, 5),
    "damage": (np.float64, 1), "ddamage": (np.float64, 1),
    "m": (np.float64, 1),
    "p": (np.float64, 1), "cs": (np.float64, 1), "reduce": (np.float64, 1),
    "divv": (np.float64, 1), "gradv": (np.float64, 6),
    "eps_min": (np.float64, 1), "m_zero": (np.float64, 1), "growth": (np.float64, 1),
    "flag": (np.uint32, 1), "mat": (np.uint32, 1), "ncnt": (np.uint32, 1), "n_flaws": (np.uint32, 1),
}


def basalt(**overrides) -> dict:
    """Default basalt body parameters (core/system/Settings.cpp:793-880)."""
    m = dict(
        rho0=2700.0, A=2.67e10, B=2.67e10, a=0.5, b=1.5, alpha=5.0, beta=5.0,
        u0=4.87e8, uiv=4.72e6, ucv=1.82e7,
        yielding=0, damage=0,
        shear_modulus=2.27e10, elasticity_limit=3.5e9, melt_energy=3.4e6, young_modulus=5.7e10,
        rho_lo=50.0, rho_hi=INFTY, u_lo=0.0, u_hi=INFTY, d_lo=0.0, d_hi=1.0,
        rho_min=100.0, u_min=1.0, s_min=1.0e5, d_min=0.03,
        cohesion=9.0e7, internal_friction=2.0, dry_friction=0.8,
    )
    m.update(overrides)
    return m


@dataclasses.dataclass
class RunConfig:
    """Term set and integrator settings (the subset of RunSettings the hot path reads)."""
    solid_stress: int = 0          # ForceEnum::SOLID_STRESS
    adaptive_h: int = 0            # AdaptiveSmoothingLength (else ConstSmoothingLength)
    sum_only_undamaged: int = 1    # SPH_SUM_ONLY_UNDAMAGED (core/system/Settings.cpp:600)
    has_reduce: int = 0            # STRESS_REDUCING present (a material has a rheology)
    has_damage: int = 0            # DAMAGE present
    av_alpha: float = 1.5
    av_beta: float = 3.0
    h_lo: float = 1e-5
    h_hi: float = 1e10
    criteria: int = 1              # bit0 COURANT, bit1 DERIVATIVES, bit2 ACCELERATION, bit3 DIVERGENCE
    courant: float = 0.2
    deriv_factor: float = 0.2
    div_factor: float = 0.005
    max_dt: float = 10.0
    initial_dt: float = 0.03
    # GravitySolver<SymmetricSolver> (core/sph/solvers/GravitySolver.cpp:64-91): self-gravity added to the
    # accelerations of every derivative evaluation; 0 none, 1 BruteForceGravity, 2 BarnesHut, both with the
    # SPH-kernel (M4) softening (RunSettingsId::GRAVITY_*, defaults core/system/Settings.cpp:633-642).
    # As in Factory::getGravity (core/system/Factory.cpp:361-413), gravity_G applies to Barnes-Hut only;
    # the direct sum always uses Constants::gravity.
    gravity: int = 0
    gravity_G: float = 6.67408e-11  # Constants::gravity (core/physics/Constants.h:32)
    gravity_theta: float = 0.5
    gravity_order: int = 3
    gravity_leaf: int = 25


class ParticleState:
    """SoA particle state in storage order plus the material table."""

    def __init__(self, n: int, materials: List[dict]):
        self.n = int(n)
        self.materials = [dict(m) for m in materials]
        self.a: Dict[str, np.ndarray] = {}
        for name, (dt, c) in FIELDS.items():
            shape = (self.n,) if c == 1 else (self.n, c)
            self.a[name] = np.zeros(shape, dtype=dt)

    def __getitem__(self, k):
        return self.a[k]

    def __setitem__(self, k, v):
        self.a[k][...] = v

    def copy(self) -> "ParticleState":
        out = ParticleState(self.n, self.materials)
        for k, v in self.a.items():
            out.a[k][...] = v
        return out

    def subset(self, idx) -> "ParticleState":
        idx = np.asarray(idx, dtype=np.int64)
        out = ParticleState(len(idx), self.materials)
        for k, v in self.a.items():
            out.a[k][...] = v[idx]
        return out

"""Seeded input states for the parity tests.

`exercised_impact` starts from the two-body impact geometry and perturbs it so that one derivative
evaluation exercises every branch of the path: bodies in contact (impactor offset -0.5 h), a
velocity field with shear and compression, densities above and below rho0 (compressed and expanded
Tillotson phases), internal energies across the incipient/complete vaporisation band (blended phase),
deviatoric stresses above the von Mises limit, damage values in [0, 1] including fully damaged
particles (D = 1, derivative set to LARGE), and positive/negative pressures.
"""
import numpy as np

from opensph_amd import initial


def exercised_impact(n_target=6000, n_projectile=600, seed=7):
    st, cfg = initial.impact(n_target=n_target, n_projectile=n_projectile, r_target=50e3, r_projectile=12e3,
                             offset_h=-0.5, seed=seed)
    rng = np.random.default_rng(seed + 100)
    n = st.n
    st["v"][:, :3] += rng.normal(0.0, 50.0, size=(n, 3))
    st["v"][:, 0] += 1e-3 * st["r"][:, 1]          # shear
    st["rho"][:] = st.materials[0]["rho0"] * rng.uniform(0.85, 1.15, size=n)
    u = rng.uniform(0.0, 3e6, size=n)
    hot = rng.random(n) < 0.2
    u[hot] = rng.uniform(4.0e6, 2.0e7, size=hot.sum())  # around uiv..ucv
    st["u"][:] = u
    st["s"][:] = rng.normal(0.0, 2.0e9, size=(n, 5))
    d = rng.uniform(0.0, 1.0, size=n) ** 3
    d[rng.random(n) < 0.05] = 1.0
    st["damage"][:] = d
    st["v"][:, 3] = rng.normal(0.0, 1.0, size=n)
    return st, cfg


def exercised_fluid(seed=3):
    st, cfg = initial.fluid_sphere(n_requested=4000, radius=1000.0, sigma_v=20.0, seed=seed)
    rng = np.random.default_rng(seed + 1)
    st["rho"][:] = st.materials[0]["rho0"] * rng.uniform(0.9, 1.1, size=st.n)
    st["u"][:] = rng.uniform(0.0, 1.0e7, size=st.n)
    return st, cfg


def exercised_impact_drucker_prager(seed=7):
    """exercised_impact with DruckerPragerRheology: pressures of both signs around the cohesion, the
    dry-friction branch (mu_d p above the intact strength at high pressure), damage, melting."""
    st, cfg = exercised_impact(seed=seed)
    st.materials = [dict(m, yielding=2) for m in st.materials]
    return st, cfg

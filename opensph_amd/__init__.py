"""MI355X-native OpenSPH per-timestep hot path (HIP for gfx950 behind a C ABI)."""
from .state import ParticleState, RunConfig, basalt  # noqa: F401

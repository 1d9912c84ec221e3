"""ctypes binding of the self-gravity C ABI (include/osph_gravity.h).

`GpuGravity` mirrors the reference's IGravity (core/gravity/IGravity.h:16-64) as created by
Factory::getGravity (core/system/Factory.cpp:361-413): `build()` is IGravity::build, `eval_self()`
IGravity::evalSelfGravity (adds to the given accelerations), `eval_points()` IGravity::evalAcceleration,
`energy()` BruteForceGravity::evalEnergy, `moments()` BarnesHut::getMoments; `leapfrog()` runs resident
LeapFrog steps with the gravity-only HardSphereSolver.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional, Tuple

import numpy as np

from .gpu import GpuError, load_library

G_SI = 6.67408e-11  # Constants::gravity (core/physics/Constants.h:32)
SOLVER = {"brute_force": 0, "barnes_hut": 1}
KERNEL = {"point": 0, "m4": 1}
ORDER = {"monopole": 0, "quadrupole": 2, "octupole": 3, 0: 0, 2: 2, 3: 3}

EXPORTS = ["osph_gravity_create", "osph_gravity_destroy", "osph_gravity_last_error", "osph_gravity_set_precision",
           "osph_gravity_tune", "osph_gravity_set_particles",
           "osph_gravity_get", "osph_gravity_build", "osph_gravity_eval_self", "osph_gravity_eval_points",
           "osph_gravity_direct", "osph_gravity_energy", "osph_gravity_moments", "osph_gravity_step_leapfrog",
           "osph_gravity_integrals", "osph_gravity_sync", "osph_gravity_profile_enable",
           "osph_gravity_profile_read"]


class GravityConfig(C.Structure):
    _fields_ = [("solver", C.c_int), ("kernel", C.c_int), ("theta", C.c_double), ("order", C.c_int),
                ("leaf_size", C.c_int), ("G", C.c_double)]


class GravityStats(C.Structure):
    _fields_ = [("nodes_approx", C.c_uint64), ("nodes_exact", C.c_uint64), ("node_count", C.c_uint64),
                ("expansions", C.c_uint64), ("pairs", C.c_uint64)]


_bound = False


def _bind(L):
    global _bound
    if _bound:
        return L
    vp, i, sz, dp = C.c_void_p, C.c_int, C.c_size_t, C.POINTER(C.c_double)
    L.osph_gravity_create.argtypes = [C.POINTER(vp), C.POINTER(GravityConfig)]
    L.osph_gravity_destroy.argtypes = [vp]
    L.osph_gravity_destroy.restype = None
    L.osph_gravity_last_error.argtypes = [vp]
    L.osph_gravity_last_error.restype = C.c_char_p
    L.osph_gravity_set_precision.argtypes = [vp, i]
    L.osph_gravity_tune.argtypes = [vp, C.c_char_p, C.c_longlong]
    L.osph_gravity_set_particles.argtypes = [vp, sz, vp, vp, sz, vp]
    L.osph_gravity_get.argtypes = [vp, i, vp, sz]
    L.osph_gravity_build.argtypes = [vp]
    L.osph_gravity_eval_self.argtypes = [vp, vp, sz, C.POINTER(GravityStats)]
    L.osph_gravity_eval_points.argtypes = [vp, vp, sz, vp, sz]
    L.osph_gravity_direct.argtypes = [vp, vp, sz, vp, sz]
    L.osph_gravity_energy.argtypes = [vp, dp]
    L.osph_gravity_moments.argtypes = [vp, dp]
    L.osph_gravity_step_leapfrog.argtypes = [vp, C.c_double, i]
    L.osph_gravity_integrals.argtypes = [vp, dp]
    L.osph_gravity_sync.argtypes = [vp]
    L.osph_gravity_profile_enable.argtypes = [vp, i]
    L.osph_gravity_profile_read.argtypes = [vp, C.c_char_p, sz, dp, C.POINTER(C.c_uint64), i, C.POINTER(i)]
    _bound = True
    return L


def _vec4(a: np.ndarray) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    out = np.zeros((a.shape[0], 4))
    out[:, :a.shape[1]] = a
    return out


class GpuGravity:
    """HBM-resident self-gravity of a particle set (positions with h in lane 3, masses, velocities)."""

    def __init__(self, r: np.ndarray, m: np.ndarray, v: Optional[np.ndarray] = None, solver: str = "barnes_hut",
                 kernel: str = "m4", theta: float = 0.5, order="quadrupole", leaf_size: int = 25, G: float = G_SI,
                 precision: str = "fp64"):
        self.L = _bind(load_library())
        cfg = GravityConfig(SOLVER[solver], KERNEL[kernel], float(theta), ORDER[order], int(leaf_size), float(G))
        self.ctx = C.c_void_p()
        rc = self.L.osph_gravity_create(C.byref(self.ctx), C.byref(cfg))
        if rc != 0:
            raise GpuError(f"osph_gravity_create failed ({rc}): invalid settings or no usable GPU")
        self.G = G
        self.precision = precision
        self._check(self.L.osph_gravity_set_precision(self.ctx, {"fp64": 0, "mixed": 1}[precision]),
                    "osph_gravity_set_precision")
        self.set_particles(r, m, v)

    def _check(self, rc, what):
        if rc != 0:
            raise GpuError(f"{what} failed ({rc}): {self.L.osph_gravity_last_error(self.ctx).decode()}")

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            self.L.osph_gravity_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, **kw):
        """Launch tuning (see osph_gravity_tune): walk_screen = -1, 0 or 1; reorder_period >= 0."""
        for k, v in kw.items():
            self._check(self.L.osph_gravity_tune(self.ctx, k.encode(), int(v)), f"osph_gravity_tune({k})")

    def set_particles(self, r, m, v=None):
        r4 = _vec4(r)
        self.n = r4.shape[0]
        m = np.ascontiguousarray(m, dtype=np.float64)
        v4 = None if v is None else _vec4(v)
        self._check(self.L.osph_gravity_set_particles(self.ctx, self.n, r4.ctypes.data,
                                                      None if v4 is None else v4.ctypes.data, 32, m.ctypes.data),
                    "osph_gravity_set_particles")

    def build(self):
        self._check(self.L.osph_gravity_build(self.ctx), "osph_gravity_build")

    def eval_self(self, dv: Optional[np.ndarray] = None) -> Tuple[np.ndarray, Dict[str, int]]:
        """dv + gravitational acceleration (dv defaults to zero); returns (n, 3) and the statistics."""
        out = _vec4(dv if dv is not None else np.zeros((self.n, 3)))
        st = GravityStats()
        self._check(self.L.osph_gravity_eval_self(self.ctx, out.ctypes.data, 32, C.byref(st)),
                    "osph_gravity_eval_self")
        return out[:, :3].copy(), {name: getattr(st, name) for name, _ in GravityStats._fields_}

    def eval_points(self, p: np.ndarray) -> np.ndarray:
        p4 = _vec4(p)
        out = np.zeros_like(p4)
        self._check(self.L.osph_gravity_eval_points(self.ctx, p4.ctypes.data, p4.shape[0], out.ctypes.data, 32),
                    "osph_gravity_eval_points")
        return out[:, :3]

    def direct(self, targets: np.ndarray) -> np.ndarray:
        t = np.ascontiguousarray(targets, dtype=np.uint32)
        out = np.zeros((t.size, 4))
        self._check(self.L.osph_gravity_direct(self.ctx, t.ctypes.data, t.size, out.ctypes.data, 32),
                    "osph_gravity_direct")
        return out[:, :3]

    def energy(self) -> float:
        e = C.c_double()
        self._check(self.L.osph_gravity_energy(self.ctx, C.byref(e)), "osph_gravity_energy")
        return e.value

    def moments(self) -> np.ndarray:
        out = np.zeros(16)
        self._check(self.L.osph_gravity_moments(self.ctx, out.ctypes.data_as(C.POINTER(C.c_double))),
                    "osph_gravity_moments")
        return out

    def leapfrog(self, dt: float, steps: int = 1):
        self._check(self.L.osph_gravity_step_leapfrog(self.ctx, float(dt), int(steps)), "osph_gravity_step_leapfrog")

    def get(self, what: str) -> np.ndarray:
        k = {"r": 0, "v": 1, "dv": 2}[what]
        out = np.zeros((self.n, 4))
        self._check(self.L.osph_gravity_get(self.ctx, k, out.ctypes.data, 32), "osph_gravity_get")
        return out if what != "dv" else out[:, :3]

    def integrals(self) -> Dict[str, object]:
        out = np.zeros(11)
        self._check(self.L.osph_gravity_integrals(self.ctx, out.ctypes.data_as(C.POINTER(C.c_double))),
                    "osph_gravity_integrals")
        return {"kinetic_energy": out[0], "momentum": out[1:4], "angular_momentum": out[4:7],
                "center_of_mass": out[7:10], "mass": out[10]}

    def sync(self):
        self._check(self.L.osph_gravity_sync(self.ctx), "osph_gravity_sync")

    def profile(self, enable: bool = True):
        self._check(self.L.osph_gravity_profile_enable(self.ctx, int(enable)), "osph_gravity_profile_enable")

    def profile_read(self) -> Dict[str, Tuple[float, int]]:
        cap, name_cap = 64, 64
        names = C.create_string_buffer(cap * name_cap)
        ms = (C.c_double * cap)()
        launches = (C.c_uint64 * cap)()
        count = C.c_int()
        self._check(self.L.osph_gravity_profile_read(self.ctx, names, name_cap, ms, launches, cap, C.byref(count)),
                    "osph_gravity_profile_read")
        out = {}
        for k in range(min(count.value, cap)):
            name = names.raw[k * name_cap:(k + 1) * name_cap].split(b"\0", 1)[0].decode()
            out[name] = (ms[k], launches[k])
        return out

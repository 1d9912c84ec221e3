"""ctypes binding of libosph_gpu.so (the C ABI in include/osph_gpu.h).

`GpuSolver` is the host-side counterpart of the reference's SymmetricSolver<3> + PredictorCorrector
pair for this path: it uploads a `ParticleState` into HBM once, then `integrate()` mirrors
ISolver::integrate (core/timestepping/ISolver.h:38) and `step()` mirrors ITimeStepping::step of
PredictorCorrector (core/timestepping/TimeStepping.cpp:34-75,286-346).  There is no CPU fallback:
if the library is missing or no GPU is visible, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Tuple

import numpy as np

from .state import ParticleState, RunConfig

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libosph_gpu.so")

# quantity ids (include/osph_gpu.h)
Q = dict(POSITION=0, MASS=1, DENSITY=2, ENERGY=3, DEVIATORIC_STRESS=4, DAMAGE=5, PRESSURE=6, SOUND_SPEED=7,
         STRESS_REDUCING=8, VELOCITY_DIVERGENCE=9, VELOCITY_GRADIENT=10, NEIGHBOR_CNT=11, FLAG=12,
         MATERIAL_ID=13, EPS_MIN=14, M_ZERO=15, EXPLICIT_GROWTH=16, N_FLAWS=17)
PRECISION = {"fp64": 0, "mixed": 1}

EXPORTS = ["osph_create", "osph_destroy", "osph_last_error", "osph_device_info", "osph_resize",
           "osph_set_quantity", "osph_get_quantity", "osph_zero_derivatives", "osph_build_neighbors",
           "osph_get_pairs", "osph_integrate", "osph_step_pc", "osph_timestep", "osph_sync", "osph_kernel_lut",
           "osph_stream_copy", "osph_tune", "osph_integrals", "osph_step_pc_batch",
           "osph_profile_enable", "osph_profile_read", "osph_adjoint_forward", "osph_adjoint_backward",
           "osph_adjoint_clear", "osph_attach_gravity", "osph_set_flaws", "osph_setup_spheres",
           "osph_integrals_noalloc", "osph_autotune"]


class Material(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("rho0", "A", "B", "a", "b", "alpha", "beta", "u0", "uiv", "ucv")] + [
        ("yielding", C.c_int), ("damage", C.c_int)] + [
        (k, C.c_double) for k in ("shear_modulus", "elasticity_limit", "melt_energy", "young_modulus",
                                  "rho_lo", "rho_hi", "u_lo", "u_hi", "d_lo", "d_hi",
                                  "rho_min", "u_min", "s_min", "d_min",
                                  "cohesion", "internal_friction", "dry_friction")]


class Config(C.Structure):
    _fields_ = [("solid_stress", C.c_int), ("adaptive_h", C.c_int), ("sum_only_undamaged", C.c_int),
                ("has_reduce", C.c_int), ("has_damage", C.c_int),
                ("av_alpha", C.c_double), ("av_beta", C.c_double), ("h_lo", C.c_double), ("h_hi", C.c_double),
                ("criteria", C.c_int), ("courant", C.c_double), ("deriv_factor", C.c_double),
                ("div_factor", C.c_double), ("max_dt", C.c_double), ("precision", C.c_int)]


class SphereBody(C.Structure):
    _fields_ = [("particle_count", C.c_double), ("radius", C.c_double), ("center", C.c_double * 3),
                ("eta", C.c_double), ("velocity", C.c_double * 3), ("material", C.c_int), ("flag", C.c_uint32)]


class Collision(C.Structure):
    _fields_ = [("impact_angle_deg", C.c_double), ("impactor_offset_h", C.c_double)]


class BodyInfo(C.Structure):
    _fields_ = [("first", C.c_size_t), ("count", C.c_size_t), ("volume", C.c_double), ("h", C.c_double),
                ("mass", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("neighbors_min", C.c_uint32), ("neighbors_max", C.c_uint32), ("neighbors_mean", C.c_double)]


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libosph_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m opensph_amd.build` (no CPU fallback)")
    L = C.CDLL(path)
    vp, i, sz, dp = C.c_void_p, C.c_int, C.c_size_t, C.POINTER(C.c_double)
    L.osph_create.argtypes = [C.POINTER(vp), C.POINTER(Config), C.POINTER(Material), i]
    L.osph_destroy.argtypes = [vp]
    L.osph_destroy.restype = None
    L.osph_last_error.argtypes = [vp]
    L.osph_last_error.restype = C.c_char_p
    L.osph_device_info.argtypes = [C.c_char_p, sz, C.POINTER(i)]
    L.osph_resize.argtypes = [vp, sz]
    L.osph_set_quantity.argtypes = [vp, i, i, vp, sz]
    L.osph_get_quantity.argtypes = [vp, i, i, vp, sz]
    L.osph_zero_derivatives.argtypes = [vp]
    L.osph_build_neighbors.argtypes = [vp]
    L.osph_get_pairs.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), sz, C.POINTER(sz)]
    L.osph_integrate.argtypes = [vp, C.POINTER(Stats)]
    L.osph_step_pc.argtypes = [vp, C.c_double, dp, C.POINTER(i)]
    L.osph_timestep.argtypes = [vp, dp, C.POINTER(i)]
    L.osph_sync.argtypes = [vp]
    L.osph_kernel_lut.argtypes = [dp, dp]
    L.osph_stream_copy.argtypes = [sz, i, dp]
    L.osph_tune.argtypes = [vp, C.c_char_p, C.c_longlong]
    L.osph_autotune.argtypes = [vp, C.c_double, i, C.POINTER(C.c_char_p), C.POINTER(C.c_int),
                                C.POINTER(C.c_longlong), i, C.POINTER(C.c_longlong), dp, dp, i, C.POINTER(i)]
    L.osph_integrals.argtypes = [vp, dp]
    L.osph_integrals_noalloc.argtypes = [vp, dp]
    L.osph_step_pc_batch.argtypes = [C.POINTER(vp), i, dp, dp, C.POINTER(i)]
    L.osph_adjoint_forward.argtypes = [vp, i, dp, i]
    L.osph_adjoint_backward.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(C.c_uint)]
    L.osph_adjoint_clear.argtypes = [vp]
    L.osph_attach_gravity.argtypes = [vp, vp]
    L.osph_set_flaws.argtypes = [vp, sz, sz, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64]
    L.osph_setup_spheres.argtypes = [vp, C.POINTER(SphereBody), i, C.POINTER(Collision), C.POINTER(BodyInfo)]
    L.osph_profile_enable.argtypes = [vp, i]
    L.osph_profile_read.argtypes = [vp, C.c_char_p, sz, dp, C.POINTER(C.c_uint64), i, C.POINTER(i)]
    _lib = L
    return L


def kernel_lut() -> Tuple[np.ndarray, np.ndarray]:
    """The M4 look-up tables the device uses (host-built, no GPU call)."""
    L = load_library()
    v = np.zeros(40001)
    g = np.zeros(40001)
    L.osph_kernel_lut(v.ctypes.data_as(C.POINTER(C.c_double)), g.ctypes.data_as(C.POINTER(C.c_double)))
    return v, g


def _material(m: dict) -> Material:
    from .state import basalt
    defaults = basalt()  # parameters a material dict leaves out take the reference's defaults
    out = Material()
    for name, t in Material._fields_:
        setattr(out, name, t(m.get(name, defaults[name])).value)
    return out


def _config(cfg: RunConfig, precision: str) -> Config:
    c = Config()
    for name, t in Config._fields_:
        if name == "precision":
            continue
        setattr(c, name, t(getattr(cfg, name)).value)
    c.precision = PRECISION[precision]
    return c


# ParticleState field -> (quantity, order, host layout builder)
def _vec4(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _traceless(a):
    out = np.zeros((a.shape[0], 8))
    out[:, :5] = a
    return out


def _symmetric(a):
    out = np.zeros((a.shape[0], 8))
    out[:, 0:3] = a[:, 0:3]
    out[:, 4:7] = a[:, 3:6]
    return out


_UPLOAD = [
    ("r", "POSITION", 0), ("v", "POSITION", 1), ("dv", "POSITION", 2),
    ("m", "MASS", 0), ("rho", "DENSITY", 0), ("drho", "DENSITY", 1), ("u", "ENERGY", 0), ("du", "ENERGY", 1),
    ("s", "DEVIATORIC_STRESS", 0), ("ds", "DEVIATORIC_STRESS", 1), ("damage", "DAMAGE", 0),
    ("ddamage", "DAMAGE", 1), ("p", "PRESSURE", 0), ("cs", "SOUND_SPEED", 0), ("reduce", "STRESS_REDUCING", 0),
    ("divv", "VELOCITY_DIVERGENCE", 0), ("gradv", "VELOCITY_GRADIENT", 0), ("ncnt", "NEIGHBOR_CNT", 0),
    ("flag", "FLAG", 0), ("mat", "MATERIAL_ID", 0), ("eps_min", "EPS_MIN", 0), ("m_zero", "M_ZERO", 0),
    ("growth", "EXPLICIT_GROWTH", 0), ("n_flaws", "N_FLAWS", 0),
]


class GpuError(RuntimeError):
    pass


class GpuSolver:
    """HBM-resident SymmetricSolver<3> + PredictorCorrector for one ParticleState."""

    def __init__(self, st: ParticleState, cfg: RunConfig, precision: str = "fp64"):
        self.L = load_library()
        self.cfg = cfg
        self.precision = precision
        self.materials = st.materials
        mats = (Material * len(st.materials))(*[_material(m) for m in st.materials])
        self._mats = mats
        self._cfg = _config(cfg, precision)
        self.ctx = C.c_void_p()
        rc = self.L.osph_create(C.byref(self.ctx), C.byref(self._cfg), mats, len(st.materials))
        if rc != 0:
            raise GpuError(f"osph_create failed ({rc}): no usable GPU or HIP runtime error")
        self.n = 0
        self.dt = cfg.initial_dt
        self.gravity = None
        self.upload(st)
        if getattr(cfg, "gravity", 0):
            self.attach_gravity_from_config()

    @classmethod
    def from_bodies(cls, bodies, materials, cfg: RunConfig, precision: str = "fp64", collision=None):
        """A solver whose state is built on the device by osph_setup_spheres: `bodies` is a list of dicts
        (particle_count, radius, center, eta, velocity, material, flag), `collision` None or (impact angle
        in degrees, impactor offset in target h).  Returns (solver, [(first, count, volume, h) per body]);
        the materials' density floors are updated as the device updated them."""
        st = ParticleState(0, materials)
        solver = cls(st, cfg, precision=precision)
        solver.materials = [dict(m) for m in materials]
        info = solver.setup_spheres(bodies, collision)
        return solver, info

    def setup_spheres(self, bodies, collision=None):
        """osph_setup_spheres on this context (see from_bodies)."""
        arr = (SphereBody * len(bodies))()
        for k, b in enumerate(bodies):
            arr[k].particle_count = float(b["particle_count"])
            arr[k].radius = float(b["radius"])
            arr[k].center = (C.c_double * 3)(*[float(x) for x in b.get("center", (0.0, 0.0, 0.0))])
            arr[k].eta = float(b.get("eta", 1.3))
            arr[k].velocity = (C.c_double * 3)(*[float(x) for x in b.get("velocity", (0.0, 0.0, 0.0))])
            arr[k].material = int(b.get("material", 0))
            arr[k].flag = int(b.get("flag", k))
        col = None
        if collision is not None:
            col = Collision(float(collision[0]), float(collision[1]))
        info = (BodyInfo * max(len(bodies), 1))()
        self._check(self.L.osph_setup_spheres(self.ctx, arr, len(bodies), C.byref(col) if col is not None else None,
                                              info), "osph_setup_spheres")
        out = [(int(info[k].first), int(info[k].count), float(info[k].volume), float(info[k].h))
               for k in range(len(bodies))]
        self.n = sum(c for _, c, _, _ in out)
        # the density floor of ContinuityEquation::create, as the device applied it to its material table
        w0 = 1.0 / 3.14159265358979323846264338327950288
        limit = min((info[k].mass * w0 / (info[k].h * info[k].h * info[k].h) for k in range(len(bodies))),
                    default=float(np.float32(1e20)))
        for m in self.materials:
            m["rho_lo"] = max(limit, m["rho_lo"])
        return out

    def attach_gravity_from_config(self):
        """GravitySolver<SymmetricSolver> with the gravity Factory::getGravity builds from the run settings
        (core/system/Factory.cpp:361-413): the direct sum with Constants::gravity, or Barnes-Hut with the
        configured opening angle, multipole order, leaf size and constant; M4 (SPH-kernel) softening."""
        from .gravity import G_SI, GpuGravity
        st = self.download()
        brute = self.cfg.gravity == 1
        g = GpuGravity(st["r"], st["m"], solver="brute_force" if brute else "barnes_hut", kernel="m4",
                       theta=self.cfg.gravity_theta, order=self.cfg.gravity_order, leaf_size=self.cfg.gravity_leaf,
                       G=G_SI if brute else self.cfg.gravity_G, precision=self.precision)
        self.attach_gravity(g)

    def set_flaws(self, first: int, count: int, volume: float, growth_speed: float, seed: int,
                  k_weibull: float = 4.0e35, m_weibull: float = 9.0):
        """osph_set_flaws: Grady-Kipp flaws of the body [first, first + count) drawn on the device."""
        self._check(self.L.osph_set_flaws(self.ctx, int(first), int(count), float(volume), float(k_weibull),
                                          float(m_weibull), float(growth_speed), int(seed) & ((1 << 64) - 1)),
                    "osph_set_flaws")

    def attach_gravity(self, g):
        """osph_attach_gravity: g (a GpuGravity, or None to detach) evaluates the self-gravity of every
        derivative evaluation."""
        self._check(self.L.osph_attach_gravity(self.ctx, g.ctx if g is not None else None), "osph_attach_gravity")
        self.gravity = g

    def _check(self, rc, what):
        if rc != 0:
            msg = self.L.osph_last_error(self.ctx).decode()
            raise GpuError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            self.L.osph_destroy(self.ctx)
            self.ctx = C.c_void_p()
        if getattr(self, "gravity", None) is not None:
            self.gravity.close()  # after the context it is attached to
            self.gravity = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, st: ParticleState):
        self._check(self.L.osph_resize(self.ctx, st.n), "osph_resize")
        self.n = st.n
        for field, q, order in _UPLOAD:
            a = st[field]
            if field == "s" or field == "ds":
                host = _traceless(a)
            elif field == "gradv":
                host = _symmetric(a)
            elif a.ndim == 2:
                host = _vec4(a)
            else:
                host = np.ascontiguousarray(a)
            self._check(self.L.osph_set_quantity(self.ctx, Q[q], order, host.ctypes.data, 0),
                        f"osph_set_quantity({field})")

    def download(self, template: ParticleState | None = None) -> ParticleState:
        st = ParticleState(self.n, self.materials) if template is None else template.copy()
        for field, q, order in _UPLOAD:
            a = st.a[field]
            if field in ("s", "ds", "gradv"):
                host = np.zeros((self.n, 8))
            elif a.ndim == 2:
                host = np.zeros((self.n, 4))
            else:
                host = np.zeros(self.n, dtype=a.dtype)
            self._check(self.L.osph_get_quantity(self.ctx, Q[q], order, host.ctypes.data, 0),
                        f"osph_get_quantity({field})")
            if field in ("s", "ds"):
                a[:] = host[:, :5]
            elif field == "gradv":
                a[:, 0:3] = host[:, 0:3]
                a[:, 3:6] = host[:, 4:7]
            elif field == "dv":
                a[:, :3] = host[:, :3]
                a[:, 3] = 0.0
            else:
                a[...] = host
        return st

    def save(self, path, run_time: float = 0.0, template: ParticleState | None = None):
        """BinaryOutput::dump of the resident state (core/io/Output.cpp:481-595): the quantities of the path,
        the materials' parameters and the current time step, in the reference's .ssf format."""
        from . import ssf
        st = self.download(template)
        ssf.write(path, ssf.from_state(st, run_time=run_time, time_step=self.dt, solid=bool(self.cfg.solid_stress),
                                       damage=bool(self.cfg.has_damage)))

    @classmethod
    def load(cls, path, cfg: RunConfig, precision: str = "fp64") -> "GpuSolver":
        """BinaryInput::load (core/io/Output.cpp:697-803) of an .ssf file into a new resident solver, which
        continues with the file's time step."""
        from . import ssf
        f = ssf.read(path)
        solver = cls(ssf.to_state(f), cfg, precision=precision)
        solver.dt = f.time_step
        return solver

    def zero_derivatives(self):
        self._check(self.L.osph_zero_derivatives(self.ctx), "osph_zero_derivatives")

    def integrate(self) -> Dict[str, float]:
        """zeroHighestDerivatives + SymmetricSolver<3>::integrate."""
        self.zero_derivatives()
        s = Stats()
        self._check(self.L.osph_integrate(self.ctx, C.byref(s)), "osph_integrate")
        return dict(neighbors_min=s.neighbors_min, neighbors_max=s.neighbors_max, neighbors_mean=s.neighbors_mean)

    def step(self, dt: float | None = None) -> Tuple[float, float, int]:
        """One PredictorCorrector step; returns (dt used, next dt, CriterionId)."""
        dt = self.dt if dt is None else dt
        nxt = C.c_double()
        crit = C.c_int()
        self._check(self.L.osph_step_pc(self.ctx, dt, C.byref(nxt), C.byref(crit)), "osph_step_pc")
        self.dt = nxt.value
        return dt, nxt.value, crit.value

    def timestep(self) -> Tuple[float, int]:
        dt = C.c_double()
        crit = C.c_int()
        self._check(self.L.osph_timestep(self.ctx, C.byref(dt), C.byref(crit)), "osph_timestep")
        return dt.value, crit.value

    def integrals(self) -> Dict[str, object]:
        """Integrals.cpp diagnostics of the resident state (osph_integrals)."""
        out = np.zeros(13)
        self._check(self.L.osph_integrals(self.ctx, out.ctypes.data_as(C.POINTER(C.c_double))), "osph_integrals")
        return {"energy": out[0], "kinetic_energy": out[1], "internal_energy": out[2], "momentum": out[3:6],
                "angular_momentum": out[6:9], "center_of_mass": out[9:12], "mass": out[12]}

    def integrals_noalloc(self) -> Dict[str, object]:
        """integrals() through osph_integrals_noalloc (partial sums in the library's static scratch)."""
        out = np.zeros(13)
        self._check(self.L.osph_integrals_noalloc(self.ctx, out.ctypes.data_as(C.POINTER(C.c_double))),
                    "osph_integrals_noalloc")
        return {"energy": out[0], "kinetic_energy": out[1], "internal_energy": out[2], "momentum": out[3:6],
                "angular_momentum": out[6:9], "center_of_mass": out[9:12], "mass": out[12]}

    def adjoint_forward(self, dt, checkpoint_every: int = 10):
        """Runs len(dt) PC steps with the given time steps, checkpointing every `checkpoint_every` steps for
        a later adjoint_backward (osph_adjoint_forward)."""
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        self._check(self.L.osph_adjoint_forward(self.ctx, int(dt.size), dt.ctypes.data_as(C.POINTER(C.c_double)),
                                                int(checkpoint_every)), "osph_adjoint_forward")

    def adjoint_backward(self, seed_r=None, seed_v=None, seed_rho=None, seed_u=None) -> Dict[str, object]:
        """Gradients of J with respect to the initial r, v, rho, u and the Tillotson (a, b, A) of each
        material, given dJ/d(final r, v, rho, u) (osph_adjoint_backward).  The resident state returns to the
        initial state."""
        n = self.n

        def arr(a, shape):
            return None if a is None else np.ascontiguousarray(np.broadcast_to(a, shape), dtype=np.float64)

        sr, sv = arr(seed_r, (n, 3)), arr(seed_v, (n, 3))
        srho, su = arr(seed_rho, (n,)), arr(seed_u, (n,))
        out = {"r": np.zeros((n, 3)), "v": np.zeros((n, 3)), "rho": np.zeros(n), "u": np.zeros(n),
               "eos": np.zeros((len(self.materials), 3))}
        clamped = C.c_uint()
        ptr = lambda a: None if a is None else a.ctypes.data
        self._check(self.L.osph_adjoint_backward(self.ctx, ptr(sr), ptr(sv), ptr(srho), ptr(su), ptr(out["r"]),
                                                 ptr(out["v"]), ptr(out["rho"]), ptr(out["u"]), ptr(out["eos"]),
                                                 C.byref(clamped)), "osph_adjoint_backward")
        out["clamped"] = int(clamped.value)
        return out

    def adjoint_clear(self):
        self._check(self.L.osph_adjoint_clear(self.ctx), "osph_adjoint_clear")

    def pairs(self) -> Tuple[np.ndarray, np.ndarray]:
        self._check(self.L.osph_build_neighbors(self.ctx), "osph_build_neighbors")
        cnt = C.c_size_t()
        self._check(self.L.osph_get_pairs(self.ctx, None, None, 0, C.byref(cnt)), "osph_get_pairs")
        pi = np.zeros(cnt.value, dtype=np.uint32)
        pj = np.zeros(cnt.value, dtype=np.uint32)
        self._check(self.L.osph_get_pairs(self.ctx, pi.ctypes.data_as(C.POINTER(C.c_uint32)),
                                          pj.ctypes.data_as(C.POINTER(C.c_uint32)), cnt.value, C.byref(cnt)),
                    "osph_get_pairs")
        return pi, pj

    def tune(self, **kw):
        """Launch shapes of the list and pair kernels (see osph_tune)."""
        for k, v in kw.items():
            self._check(self.L.osph_tune(self.ctx, k.encode(), int(v)), f"osph_tune({k})")

    def sync(self):
        self._check(self.L.osph_sync(self.ctx), "osph_sync")

    def profile(self, enable=True):
        """True / 1: time every launch; 2: the pair-sum launches only; False / 0: off."""
        self._check(self.L.osph_profile_enable(self.ctx, int(enable)), "osph_profile_enable")

    def profile_read(self) -> Dict[str, Tuple[float, int]]:
        cap, name_cap = 64, 64
        names = C.create_string_buffer(cap * name_cap)
        ms = np.zeros(cap)
        launches = np.zeros(cap, dtype=np.uint64)
        cnt = C.c_int()
        self._check(self.L.osph_profile_read(self.ctx, names, name_cap, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                             launches.ctypes.data_as(C.POINTER(C.c_uint64)), cap, C.byref(cnt)),
                    "osph_profile_read")
        out = {}
        raw = names.raw
        for k in range(min(cnt.value, cap)):
            name = raw[k * name_cap:(k + 1) * name_cap].split(b"\0", 1)[0].decode()
            out[name] = (float(ms[k]), int(launches[k]))
        return out


def stream_copy(nbytes: int = 1 << 30, iters: int = 20) -> float:
    """Read+write GB/s of a plain device copy of nbytes (HBM roof and counter calibration)."""
    L = load_library()
    gbs = C.c_double()
    rc = L.osph_stream_copy(nbytes, iters, C.byref(gbs))
    if rc != 0:
        raise GpuError("osph_stream_copy failed")
    return gbs.value


def device_info() -> Tuple[str, int]:
    L = load_library()
    buf = C.create_string_buffer(256)
    cu = C.c_int()
    rc = L.osph_device_info(buf, 256, C.byref(cu))
    if rc != 0:
        raise GpuError("no HIP device")
    return buf.value.decode(), cu.value


def step_batch(solvers) -> Tuple[np.ndarray, np.ndarray]:
    """One predictor-corrector step of every solver in the list (osph_step_pc_batch), each with its own
    current time step; returns the next time steps and criterion ids and advances each solver's dt."""
    if not solvers:
        return np.zeros(0), np.zeros(0, dtype=np.int32)
    L = solvers[0].L
    n = len(solvers)
    ctxs = (C.c_void_p * n)(*[s.ctx.value for s in solvers])
    dt = np.ascontiguousarray([s.dt for s in solvers], dtype=np.float64)
    nxt = np.zeros(n)
    crit = np.zeros(n, dtype=np.int32)
    rc = L.osph_step_pc_batch(ctxs, n, dt.ctypes.data_as(C.POINTER(C.c_double)),
                              nxt.ctypes.data_as(C.POINTER(C.c_double)), crit.ctypes.data_as(C.POINTER(C.c_int)))
    if rc != 0:
        msgs = "; ".join(L.osph_last_error(s.ctx).decode() for s in solvers if L.osph_last_error(s.ctx))
        raise GpuError(f"osph_step_pc_batch failed ({rc}): {msgs}")
    for s, d in zip(solvers, nxt):
        s.dt = float(d)
    return nxt, crit


"""TEST INFRASTRUCTURE -- ctypes bridge to the CPU checkers built by oracle/Makefile.

* `Oracle`  wraps liboracle.so, the plain-C restatement of the hot path (osph_oracle.c).
* `RefRun`  wraps _ref/libosph_ref.so, the real OpenSPH core plus ref_driver.cpp (only where it was built).
* `AdapterRun` wraps _ref/libosph_adapter.so: the OpenSPH-side adapter (opensph_amd/adapter) stepping the
  reference-built Storage on the GPU, with adapter_driver.cpp.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libosph_ref.so")
ADAPTER_SO = os.path.join(HERE, "_ref", "libosph_adapter.so")

# plane indices, in the order of the enums in osph_oracle.h
_DPLANES = ["X", "Y", "Z", "H", "VX", "VY", "VZ", "VH", "AX", "AY", "AZ", "RHO", "DRHO", "U", "DU",
            "SXX", "SYY", "SXY", "SXZ", "SYZ", "DSXX", "DSYY", "DSXY", "DSXZ", "DSYZ", "D", "DD",
            "M", "P", "CS", "RED", "DIVV", "GXX", "GYY", "GZZ", "GXY", "GXZ", "GYZ",
            "EPSMIN", "MZERO", "GROWTH",
            "PAX", "PAY", "PAZ", "PDRHO", "PDU", "PDSXX", "PDSYY", "PDSXY", "PDSXZ", "PDSYZ", "PDD"]
D = {k: i for i, k in enumerate(_DPLANES)}
NP = len(_DPLANES)
_IPLANES = ["FLAG", "MAT", "NCNT", "NFLAWS"]
IP = {k: i for i, k in enumerate(_IPLANES)}
NI = len(_IPLANES)

# ParticleState field -> list of planes
_MAP = {
    "r": ["X", "Y", "Z", "H"], "v": ["VX", "VY", "VZ", "VH"], "dv": ["AX", "AY", "AZ", None],
    "rho": ["RHO"], "drho": ["DRHO"], "u": ["U"], "du": ["DU"],
    "s": ["SXX", "SYY", "SXY", "SXZ", "SYZ"], "ds": ["DSXX", "DSYY", "DSXY", "DSXZ", "DSYZ"],
    "damage": ["D"], "ddamage": ["DD"], "m": ["M"], "p": ["P"], "cs": ["CS"], "reduce": ["RED"],
    "divv": ["DIVV"], "gradv": ["GXX", "GYY", "GZZ", "GXY", "GXZ", "GYZ"],
    "eps_min": ["EPSMIN"], "m_zero": ["MZERO"], "growth": ["GROWTH"],
}
_IMAP = {"flag": "FLAG", "mat": "MAT", "ncnt": "NCNT", "n_flaws": "NFLAWS"}


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
                ("div_factor", C.c_double), ("max_dt", C.c_double),
                ("n_mat", C.c_int), ("mat", C.POINTER(Material)),
                ("gravity", C.c_int)]


def make_config(cfg, materials):
    mats = (Material * len(materials))()
    from opensph_amd.state import basalt
    defaults = basalt()  # parameters a material dict leaves out take the reference's defaults
    for i, m in enumerate(materials):
        for name, _ in Material._fields_:
            setattr(mats[i], name, type(getattr(mats[i], name))(m.get(name, defaults[name])))
    c = Config()
    for name, _ in Config._fields_:
        if name in ("n_mat", "mat"):
            continue
        setattr(c, name, getattr(cfg, name))
    c.n_mat = len(materials)
    c.mat = C.cast(mats, C.POINTER(Material))
    c._keep = mats
    return c


def to_planes(st):
    P = np.zeros((NP, st.n))
    I = np.zeros((NI, st.n), dtype=np.uint32)
    for f, planes in _MAP.items():
        a = st[f]
        for k, pl in enumerate(planes):
            if pl is None:
                continue
            P[D[pl]] = a if a.ndim == 1 else a[:, k]
    for f, pl in _IMAP.items():
        I[IP[pl]] = st[f]
    return P, I


def from_planes(st, P, I):
    for f, planes in _MAP.items():
        a = st.a[f]
        for k, pl in enumerate(planes):
            if pl is None:
                continue
            if a.ndim == 1:
                a[:] = P[D[pl]]
            else:
                a[:, k] = P[D[pl]]
    for f, pl in _IMAP.items():
        st.a[f][:] = I[IP[pl]]


def _ptr(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run make -C oracle)")
        L = self.lib = C.CDLL(path)
        L.orc_lut.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_tillotson.argtypes = [C.POINTER(Material), C.c_double, C.c_double, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
        L.orc_eigenvalues.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_find_pairs.restype = C.c_size_t
        L.orc_find_pairs.argtypes = [C.c_size_t] + [C.POINTER(C.c_double)] * 4 + [C.POINTER(C.c_uint32)] * 2 + [
            C.c_size_t]
        L.orc_integrate.argtypes = [C.POINTER(Config), C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_size_t]
        L.orc_zero_derivatives.argtypes = [C.POINTER(C.c_double), C.c_size_t]
        L.orc_step_pc.restype = C.c_double
        L.orc_step_pc.argtypes = [C.POINTER(Config), C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_size_t,
                                  C.c_double, C.POINTER(C.c_int)]
        L.orc_timestep.restype = C.c_double
        L.orc_timestep.argtypes = [C.POINTER(Config), C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_size_t,
                                   C.POINTER(C.c_int)]

    def lut(self):
        v = np.zeros(40001)
        g = np.zeros(40001)
        self.lib.orc_lut(_ptr(v), _ptr(g))
        return v, g

    def tillotson(self, material, rho, u):
        m = make_config(type("c", (), {k: 0 for k, _ in Config._fields_})(), [material]).mat
        p, cs = C.c_double(), C.c_double()
        self.lib.orc_tillotson(m, rho, u, C.byref(p), C.byref(cs))
        return p.value, cs.value

    def eigenvalues(self, t6):
        t = np.ascontiguousarray(t6, dtype=np.float64)
        out = np.zeros(3)
        self.lib.orc_eigenvalues(_ptr(t), _ptr(out))
        return out

    def pairs(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        n = r.shape[0]
        cols = [np.ascontiguousarray(r[:, k]) for k in range(4)]
        cap = max(64 * n, 1024)
        while True:
            pi = np.zeros(cap, dtype=np.uint32)
            pj = np.zeros(cap, dtype=np.uint32)
            cnt = self.lib.orc_find_pairs(n, *[_ptr(c) for c in cols], _ptr(pi, C.c_uint32), _ptr(pj, C.c_uint32),
                                          cap)
            if cnt <= cap:
                return pi[:cnt], pj[:cnt]
            cap = cnt

    def integrate(self, st, cfg):
        P, I = to_planes(st)
        c = make_config(cfg, st.materials)
        self.lib.orc_zero_derivatives(_ptr(P), st.n)
        self.lib.orc_integrate(C.byref(c), _ptr(P), _ptr(I, C.c_uint32), st.n)
        out = st.copy()
        from_planes(out, P, I)
        return out

    def run_pc(self, st, cfg, steps, dt0=None):
        """Runs `steps` PredictorCorrector steps; returns (state, list of (dt_used, next_dt, criterion))."""
        P, I = to_planes(st)
        c = make_config(cfg, st.materials)
        dt = cfg.initial_dt if dt0 is None else dt0
        log = []
        crit = C.c_int(0)
        for _ in range(steps):
            nxt = self.lib.orc_step_pc(C.byref(c), _ptr(P), _ptr(I, C.c_uint32), st.n, dt, C.byref(crit))
            log.append((dt, nxt, crit.value))
            dt = nxt
        out = st.copy()
        from_planes(out, P, I)
        return out, log


    def run_pc_fixed(self, st, cfg, dts):
        """PredictorCorrector steps with the given time steps (the criterion's proposals ignored)."""
        P, I = to_planes(st)
        c = make_config(cfg, st.materials)
        crit = C.c_int(0)
        for dt in dts:
            self.lib.orc_step_pc(C.byref(c), _ptr(P), _ptr(I, C.c_uint32), st.n, float(dt), C.byref(crit))
        out = st.copy()
        from_planes(out, P, I)
        return out


class RefRun:
    """The real OpenSPH SymmetricSolver<3> (or, with asymmetric=True, AsymmetricSolver) + PredictorCorrector
    on a copy of the given state."""

    _lib = None

    @classmethod
    def available(cls):
        return os.path.exists(REF_SO)

    @classmethod
    def lib(cls):
        if cls._lib is None:
            L = C.CDLL(REF_SO)
            L.ref_create.restype = C.c_void_p
            L.ref_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_size_t,
                                     C.c_double, C.c_int]
            L.ref_create_solver.restype = C.c_void_p
            L.ref_create_solver.argtypes = [C.POINTER(Config), C.POINTER(C.c_double), C.POINTER(C.c_uint32),
                                            C.c_size_t, C.c_double, C.c_int, C.c_int]
            L.ref_destroy.argtypes = [C.c_void_p]
            L.ref_integrate.argtypes = [C.c_void_p]
            L.ref_step.restype = C.c_double
            L.ref_step.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
            L.ref_get_dt.restype = C.c_double
            L.ref_get_dt.argtypes = [C.c_void_p]
            L.ref_read.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_size_t]
            L.ref_pairs.restype = C.c_size_t
            L.ref_pairs.argtypes = [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_size_t]
            dp, i = C.POINTER(C.c_double), C.c_int
            L.ref_gravity.argtypes = [i, i, C.c_double, i, i, C.c_double, C.c_size_t, dp, dp, dp, dp]
            L.ref_gravity_points.argtypes = [i, i, C.c_double, i, i, C.c_double, C.c_size_t, dp, dp, dp, C.c_size_t,
                                             dp]
            cls._lib = L
        return cls._lib

    def __init__(self, st, cfg, nthreads=0, dt0=None, asymmetric=False):
        L = self.lib()
        self.st = st.copy()
        self.P, self.I = to_planes(st)
        self.c = make_config(cfg, st.materials)
        dt = cfg.initial_dt if dt0 is None else dt0
        self.h = L.ref_create_solver(C.byref(self.c), _ptr(self.P), _ptr(self.I, C.c_uint32), st.n, dt, nthreads,
                                     int(asymmetric))

    def close(self):
        if self.h:
            self.lib().ref_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def integrate(self):
        self.lib().ref_integrate(self.h)
        return self.read()

    def step(self):
        crit = C.c_int(0)
        dt = self.lib().ref_get_dt(self.h)
        nxt = self.lib().ref_step(self.h, C.byref(crit))
        return dt, nxt, crit.value

    def read(self):
        self.lib().ref_read(self.h, _ptr(self.P), _ptr(self.I, C.c_uint32), self.st.n)
        out = self.st.copy()
        from_planes(out, self.P, self.I)
        return out

    def pairs(self):
        n = self.st.n
        cap = max(64 * n, 1024)
        while True:
            pi = np.zeros(cap, dtype=np.uint32)
            pj = np.zeros(cap, dtype=np.uint32)
            cnt = self.lib().ref_pairs(self.h, _ptr(pi, C.c_uint32), _ptr(pj, C.c_uint32), cap)
            if cnt <= cap:
                return pi[:cnt], pj[:cnt]
            cap = cnt


class AdapterRun:
    """GpuPredictorCorrector + GpuSymmetricSolver (the OpenSPH-side adapter) on the Storage the real OpenSPH
    core builds from a copy of the given state; read() brings the resident state back through the storage."""

    _lib = None

    @classmethod
    def available(cls):
        return os.path.exists(ADAPTER_SO)

    @classmethod
    def lib(cls):
        if cls._lib is None:
            L = C.CDLL(ADAPTER_SO)
            L.adapter_create.restype = C.c_void_p
            L.adapter_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_size_t,
                                         C.c_double, C.c_int]
            L.adapter_destroy.argtypes = [C.c_void_p]
            L.adapter_error.restype = C.c_char_p
            L.adapter_error.argtypes = [C.c_void_p]
            dp = C.POINTER(C.c_double)
            L.adapter_step.argtypes = [C.c_void_p, dp, dp, C.POINTER(C.c_int)]
            L.adapter_integrate.argtypes = [C.c_void_p]
            L.adapter_read.argtypes = [C.c_void_p, dp, C.POINTER(C.c_uint32), C.c_size_t]
            L.adapter_save.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
            L.adjoint_run.argtypes = [C.POINTER(Config), dp, C.POINTER(C.c_uint32), C.c_size_t, C.c_int, dp, C.c_int,
                                      C.c_int, dp, dp, C.POINTER(C.c_uint32), dp, dp, dp, dp, dp, C.c_size_t,
                                      C.POINTER(C.c_size_t), C.POINTER(C.c_uint), C.c_char_p, C.c_size_t]
            cls._lib = L
        return cls._lib

    def __init__(self, st, cfg, precision="fp64", dt0=None):
        L = self.lib()
        self.st = st.copy()
        self.P, self.I = to_planes(st)
        self.c = make_config(cfg, st.materials)
        dt = cfg.initial_dt if dt0 is None else dt0
        self.h = L.adapter_create(C.byref(self.c), _ptr(self.P), _ptr(self.I, C.c_uint32), st.n, dt,
                                  1 if precision == "mixed" else 0)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib().adapter_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib().adapter_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self):
        dt, nxt, crit = C.c_double(), C.c_double(), C.c_int()
        self._check(self.lib().adapter_step(self.h, C.byref(dt), C.byref(nxt), C.byref(crit)))
        return dt.value, nxt.value, crit.value

    def integrate(self):
        self._check(self.lib().adapter_integrate(self.h))
        return self.read()

    @classmethod
    def adjoint(cls, st, cfg, dts, every, seed_v=None, precision="fp64"):
        """GpuAdjointRun (the adapter's C++ forward/backward wrapper) on the Storage OpenSPH builds from st:
        returns (final state, gradients dict like GpuSolver.adjoint_backward).  The storage holds one material
        per body (run of equal flags), so out["eos"] has one row per body."""
        L = cls.lib()
        P0, I0 = to_planes(st)
        P, I = to_planes(st)
        c = make_config(cfg, st.materials)
        n = st.n
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        sv = None if seed_v is None else np.ascontiguousarray(seed_v, dtype=np.float64)
        flag = np.asarray(st["flag"])
        nbodies = 1 + int(np.count_nonzero(np.diff(flag))) if n else 0
        eos = np.zeros(3 * max(nbodies, len(st.materials)))
        out = {"r": np.zeros((n, 3)), "v": np.zeros((n, 3)), "rho": np.zeros(n), "u": np.zeros(n)}
        clamped = C.c_uint()
        count = C.c_size_t()
        err = C.create_string_buffer(512)
        rc = L.adjoint_run(C.byref(c), _ptr(P0), _ptr(I0, C.c_uint32), n, 1 if precision == "mixed" else 0,
                           _ptr(dts), int(dts.size), int(every), None if sv is None else _ptr(sv), _ptr(P),
                           _ptr(I, C.c_uint32), _ptr(out["r"]), _ptr(out["v"]), _ptr(out["rho"]), _ptr(out["u"]),
                           _ptr(eos), eos.size, C.byref(count), C.byref(clamped), err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        out["eos"] = eos[:count.value].reshape(-1, 3)
        fin = st.copy()
        from_planes(fin, P, I)
        out["clamped"] = int(clamped.value)
        return fin, out

    def save(self, path, run_time=0.0):
        """The resident state through the storage into the reference's BinaryOutput (an .ssf file)."""
        self._check(self.lib().adapter_save(self.h, str(path).encode(), float(run_time)))

    def read(self):
        self._check(self.lib().adapter_read(self.h, _ptr(self.P), _ptr(self.I, C.c_uint32), self.st.n))
        out = self.st.copy()
        from_planes(out, self.P, self.I)
        return out


class GravityOracle:
    """The C restatement of the reference's self-gravity (osph_oracle_gravity.c)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run make -C oracle)")
        L = self.lib = C.CDLL(path)
        dp, sz = C.POINTER(C.c_double), C.c_size_t
        L.orc_gravity_lut.argtypes = [dp, dp]
        L.orc_gravity_direct.argtypes = [sz] + [dp] * 6 + [C.c_double, C.POINTER(C.c_uint32), sz, dp]
        L.orc_gravity_points.argtypes = [sz] + [dp] * 6 + [C.c_double, dp, sz, dp]
        L.orc_gravity_energy.restype = C.c_double
        L.orc_gravity_energy.argtypes = [sz] + [dp] * 6 + [C.c_double]
        L.orc_gravity_moments.argtypes = [sz] + [dp] * 4 + [dp]
        L.orc_multipole_accel.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.c_int, dp]
        L.orc_leapfrog.argtypes = [sz, dp, dp, dp, dp, dp, C.c_double, C.c_double, C.c_int]
        self.values = np.zeros(40001)
        self.grads = np.zeros(40001)
        L.orc_gravity_lut(_ptr(self.values), _ptr(self.grads))

    def _planes(self, r):
        r = np.asarray(r, dtype=np.float64)
        return [np.ascontiguousarray(r[:, k]) for k in range(4)]

    def direct(self, r, m, G, kernel="m4", targets=None):
        x, y, z, h = self._planes(r)
        m = np.ascontiguousarray(m, dtype=np.float64)
        n = len(m)
        t = None if targets is None else np.ascontiguousarray(targets, dtype=np.uint32)
        nt = n if t is None else t.size
        a = np.zeros(3 * nt)
        g = _ptr(self.grads) if kernel == "m4" else None
        self.lib.orc_gravity_direct(n, _ptr(x), _ptr(y), _ptr(z), _ptr(h), _ptr(m), g, G,
                                    None if t is None else _ptr(t, C.c_uint32), nt, _ptr(a))
        return a.reshape(3, nt).T.copy()

    def points(self, r, m, G, p, kernel="m4"):
        x, y, z, h = self._planes(r)
        m = np.ascontiguousarray(m, dtype=np.float64)
        p = np.asarray(p, dtype=np.float64)
        pp = np.ascontiguousarray(p[:, :3].T).ravel()
        a = np.zeros(3 * p.shape[0])
        g = _ptr(self.grads) if kernel == "m4" else None
        self.lib.orc_gravity_points(len(m), _ptr(x), _ptr(y), _ptr(z), _ptr(h), _ptr(m), g, G, _ptr(pp), p.shape[0],
                                    _ptr(a))
        return a.reshape(3, -1).T.copy()

    def energy(self, r, m, G, kernel="m4"):
        x, y, z, h = self._planes(r)
        m = np.ascontiguousarray(m, dtype=np.float64)
        v = _ptr(self.values) if kernel == "m4" else None
        return self.lib.orc_gravity_energy(len(m), _ptr(x), _ptr(y), _ptr(z), _ptr(h), _ptr(m), v, G)

    def moments(self, r, m):
        x, y, z, _ = self._planes(r)
        m = np.ascontiguousarray(m, dtype=np.float64)
        out = np.zeros(16)
        self.lib.orc_gravity_moments(len(m), _ptr(x), _ptr(y), _ptr(z), _ptr(m), _ptr(out))
        return out

    def multipole(self, mom13, dr, order):
        mom = np.ascontiguousarray(mom13, dtype=np.float64)
        out = np.zeros(3)
        self.lib.orc_multipole_accel(_ptr(mom), float(dr[0]), float(dr[1]), float(dr[2]), int(order), _ptr(out))
        return out

    def leapfrog(self, r, v, m, G, dt, steps, kernel="m4"):
        r = np.ascontiguousarray(np.asarray(r, dtype=np.float64).T).copy()
        v = np.ascontiguousarray(np.asarray(v, dtype=np.float64).T).copy()
        n = r.shape[1]
        a = np.zeros(3 * n)
        m = np.ascontiguousarray(m, dtype=np.float64)
        g = _ptr(self.grads) if kernel == "m4" else None
        self.lib.orc_leapfrog(n, _ptr(r), _ptr(v), _ptr(a), _ptr(m), g, G, dt, steps)
        return r.T.copy(), v.T.copy(), a.reshape(3, n).T.copy()


def ref_gravity(r, m, G, solver="barnes_hut", kernel="m4", theta=0.5, order=2, leaf=25, points=None):
    """The reference's IGravity (BruteForceGravity or BarnesHut) on (r, m): returns the self-gravity
    accelerations (n x 3) and the root moments { M, q2[5], q3[7] } (Barnes-Hut), or the accelerations at
    `points` when given."""
    L = RefRun.lib()
    r4 = np.zeros((len(m), 4))
    r = np.asarray(r, dtype=np.float64)
    r4[:, :r.shape[1]] = r
    m = np.ascontiguousarray(m, dtype=np.float64)
    sol = {"brute_force": 0, "barnes_hut": 1}[solver]
    ker = {"point": 0, "m4": 1}[kernel]
    if points is not None:
        p = np.ascontiguousarray(np.asarray(points, dtype=np.float64)[:, :3])
        a = np.zeros((p.shape[0], 3))
        L.ref_gravity_points(sol, ker, theta, order, leaf, G, len(m), _ptr(r4), _ptr(m), _ptr(p), p.shape[0], _ptr(a))
        return a
    a = np.zeros((len(m), 3))
    mom = np.zeros(13)
    L.ref_gravity(sol, ker, theta, order, leaf, G, len(m), _ptr(r4), _ptr(m), _ptr(a), _ptr(mom))
    return a, mom

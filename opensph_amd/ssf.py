"""Binary state files (.ssf) of the reference, read and written on the host.

Layout (BinaryOutput::dump, core/io/Output.cpp:481-595; BinaryInput::load, core/io/Output.cpp:697-803;
Serializer<true>, core/io/Serializer.h): little-endian; integers and enums as int64, floats as double,
fixed tags as their C arrays including the NUL ("SPH", "MAT", "NOMAT"), strings NUL-terminated.

* Header (256 bytes): "SPH", run time, particle count, quantity count (MATERIAL_ID excluded), material
  count, time step, version (BinaryIoVersion, core/io/Output.h:240-247), run type (char[16]), build date
  (char[16]), wallclock time, attractor count, 156 zero bytes.
* Quantity table: (QuantityId, OrderEnum, ValueEnum) per quantity, in the storage's id order.
* Per material (or one "NOMAT" block): "MAT", index, parameter count, parameters (BodySettingsId, type
  index of Settings' variant, value), then (QuantityId, range lower, range upper, minimal) per
  quantity; the particle range [first, last); then every quantity's buffers: values, then the first
  and second derivatives as its order has them.
* Attractors (version 20210808 on): position, velocity (Vectors), radius, mass, parameters.

A file is kept as read (header fields, the raw run-type and date fields, parameters with their type
indices and raw values), so `write(read(p))` reproduces the file byte for byte.  `to_state` and
`from_state` convert between a file and this package's ParticleState for the quantities of the
SymmetricSolver / PredictorCorrector path.  No reference counterpart beyond the format: the reference
reads and writes these files in Output.cpp.
"""
from __future__ import annotations

import dataclasses
import struct
from typing import Dict, List, Optional, Tuple

import numpy as np

from .state import INFTY, ParticleState, basalt

LATEST = 20210808  # BinaryIoVersion::LATEST
VERSIONS = (0, 20180407, 20181024, 20210320, 20210808)
PADDING = 156      # BinaryOutput::PADDING_SIZE

# QuantityId (core/quantities/QuantityIds.h)
POSITION, MASS, PRESSURE, DENSITY, ENERGY, SOUND_SPEED, DEVIATORIC_STRESS = 0, 1, 2, 3, 4, 5, 6
DAMAGE, EPS_MIN, M_ZERO, EXPLICIT_GROWTH, N_FLAWS, STRESS_REDUCING = 10, 11, 12, 13, 14, 16
VELOCITY_GRADIENT, VELOCITY_DIVERGENCE, NEIGHBOR_CNT, FLAG, MATERIAL_ID = 26, 27, 38, 42, 43

# ValueEnum (core/quantities/QuantityHelpers.h:13): components and element dtype of a buffer
SCALAR, VECTOR, TENSOR, SYMMETRIC_TENSOR, TRACELESS_TENSOR, INDEX = 0, 1, 2, 3, 4, 5
VALUE_COMPONENTS = {SCALAR: 1, VECTOR: 4, TENSOR: 9, SYMMETRIC_TENSOR: 6, TRACELESS_TENSOR: 5, INDEX: 1}

# type indices of Settings' Variant<bool, int, Float, Interval, String, Vector, SymmetricTensor,
# TracelessTensor, EnumWrapper> (core/system/Settings.h:138) -> (int64 count, double count); 4 = string
T_BOOL, T_INT, T_FLOAT, T_INTERVAL, T_STRING, T_VECTOR, T_SYMMETRIC, T_TRACELESS, T_ENUM = range(9)
PARAM_LAYOUT = {T_BOOL: ("q", 1), T_INT: ("q", 1), T_FLOAT: ("d", 1), T_INTERVAL: ("d", 2), T_VECTOR: ("d", 4),
                T_SYMMETRIC: ("d", 6), T_TRACELESS: ("d", 5), T_ENUM: ("q", 2)}

# BodySettingsId (core/system/Settings.h) of the parameters this package's materials hold
B_EOS, B_DENSITY, B_DENSITY_RANGE, B_DENSITY_MIN = 0, 7, 8, 9
B_ENERGY_RANGE, B_ENERGY_MIN, B_STRESS_TENSOR_MIN, B_DAMAGE_RANGE, B_DAMAGE_MIN = 11, 12, 14, 16, 17
B_BULK_MODULUS, B_TILLOTSON_B, B_TILLOTSON_A_SMALL, B_TILLOTSON_B_SMALL = 21, 22, 23, 24
B_TILLOTSON_ALPHA, B_TILLOTSON_BETA, B_TILLOTSON_U0, B_TILLOTSON_UIV, B_TILLOTSON_UCV = 25, 26, 27, 28, 29
B_YIELDING, B_DAMAGE_MODEL, B_SHEAR_MODULUS, B_YOUNG_MODULUS = 33, 34, 35, 36
B_ELASTICITY_LIMIT, B_MELT_ENERGY, B_COHESION, B_INTERNAL_FRICTION, B_DRY_FRICTION = 38, 39, 40, 41, 42
EOS_TILLOTSON = 4                                   # EosEnum::TILLOTSON
YIELDING_NONE, YIELDING_VON_MISES, YIELDING_DRUCKER_PRAGER = 0, 2, 3  # YieldingEnum
FRACTURE_NONE, FRACTURE_SCALAR_GRADY_KIPP = 0, 1    # FractureEnum

FLOAT_PARAMS = {B_DENSITY: "rho0", B_BULK_MODULUS: "A", B_TILLOTSON_B: "B", B_TILLOTSON_A_SMALL: "a",
                B_TILLOTSON_B_SMALL: "b", B_TILLOTSON_ALPHA: "alpha", B_TILLOTSON_BETA: "beta",
                B_TILLOTSON_U0: "u0", B_TILLOTSON_UIV: "uiv", B_TILLOTSON_UCV: "ucv",
                B_SHEAR_MODULUS: "shear_modulus", B_YOUNG_MODULUS: "young_modulus",
                B_ELASTICITY_LIMIT: "elasticity_limit", B_MELT_ENERGY: "melt_energy",
                B_DENSITY_MIN: "rho_min", B_ENERGY_MIN: "u_min", B_STRESS_TENSOR_MIN: "s_min",
                B_DAMAGE_MIN: "d_min", B_COHESION: "cohesion", B_INTERNAL_FRICTION: "internal_friction",
                B_DRY_FRICTION: "dry_friction"}
RANGE_PARAMS = {B_DENSITY_RANGE: ("rho_lo", "rho_hi"), B_ENERGY_RANGE: ("u_lo", "u_hi"),
                B_DAMAGE_RANGE: ("d_lo", "d_hi")}


class SsfError(ValueError):
    pass


@dataclasses.dataclass
class Param:
    id: int
    type: int
    value: object  # tuple of numbers, or bytes (a string without its NUL)


@dataclasses.dataclass
class Material:
    params: List[Param]
    ranges: List[Tuple[int, float, float, float]]  # (QuantityId, lower, upper, minimal)


@dataclasses.dataclass
class Attractor:
    position: Tuple[float, ...]
    velocity: Tuple[float, ...]
    radius: float
    mass: float
    params: List[Param]


@dataclasses.dataclass
class Block:
    """The particles of one material (or of the NOMAT block): [first, last) and their buffers."""
    first: int
    last: int
    buffers: Dict[int, List[np.ndarray]]  # QuantityId -> [value, dt, d2t] as its order has them


@dataclasses.dataclass
class SsfFile:
    run_time: float
    time_step: float
    version: int
    run_type: bytes        # the 16-byte field as stored
    build_date: bytes      # likewise
    wallclock: int
    quantities: List[Tuple[int, int, int]]  # (QuantityId, OrderEnum, ValueEnum)
    materials: List[Material]
    blocks: List[Block]
    attractors: List[Attractor]

    @property
    def particle_count(self) -> int:
        return sum(b.last - b.first for b in self.blocks)

    def run_type_name(self) -> str:
        return self.run_type.split(b"\0", 1)[0].decode("ascii")


class _Reader:
    def __init__(self, data: bytes):
        self.b = data
        self.d = memoryview(data)
        self.o = 0

    def take(self, n: int) -> memoryview:
        if self.o + n > len(self.d):
            raise SsfError("unexpected end of file")
        out = self.d[self.o:self.o + n]
        self.o += n
        return out

    def q(self, count=1):
        v = struct.unpack_from(f"<{count}q", self.take(8 * count))
        return v if count > 1 else v[0]

    def f(self, count=1):
        v = struct.unpack_from(f"<{count}d", self.take(8 * count))
        return v if count > 1 else v[0]

    def cstr(self) -> bytes:
        end = self.b.find(b"\0", self.o)
        if end < 0:
            raise SsfError("unterminated string")
        s = self.b[self.o:end]
        self.o = end + 1
        return s

    def array(self, dtype, count: int) -> np.ndarray:
        n = np.dtype(dtype).itemsize * count
        return np.frombuffer(self.take(n), dtype=dtype, count=count).copy()


def _read_param(r: _Reader) -> Param:
    pid, ptype = r.q(2)
    if ptype == T_STRING:
        return Param(pid, ptype, r.cstr())
    if ptype not in PARAM_LAYOUT:
        raise SsfError(f"unknown parameter type index {ptype} (parameter {pid})")
    kind, count = PARAM_LAYOUT[ptype]
    vals = r.q(count) if kind == "q" else r.f(count)
    return Param(pid, ptype, vals if count > 1 else (vals,))


def _write_param(out: List[bytes], p: Param):
    out.append(struct.pack("<2q", p.id, p.type))
    if p.type == T_STRING:
        out.append(bytes(p.value) + b"\0")
        return
    kind, count = PARAM_LAYOUT[p.type]
    out.append(struct.pack(f"<{count}{kind}", *p.value))


def read(path) -> SsfFile:
    """BinaryInput::load of a file, kept as read."""
    with open(path, "rb") as fh:
        r = _Reader(fh.read())
    if r.cstr() != b"SPH":
        raise SsfError("invalid format specifier: expected SPH")
    run_time = r.f()
    n, nq, nmat = r.q(3)
    time_step = r.f()
    version = r.q()
    if version not in VERSIONS:
        raise SsfError(f"unknown format version {version}")
    run_type = bytes(r.take(16))
    build_date = bytes(r.take(16))
    wallclock, nattr = r.q(2)
    if version < 20210808:
        nattr = 0
    r.take(PADDING)
    quantities = [tuple(r.q(3)) for _ in range(nq)]
    materials, blocks = [], []
    for mi in range(max(nmat, 1)):
        if nmat > 0:
            if r.cstr() != b"MAT":
                raise SsfError("invalid material identifier, expected MAT")
            if r.q() != mi:
                raise SsfError("unexpected material index")
            params = [_read_param(r) for _ in range(r.q())]
            ranges = []
            for qid, _, _ in quantities:
                rid = r.q()
                if rid != qid:
                    raise SsfError(f"unexpected quantity id {rid} in the ranges, expected {qid}")
                ranges.append((rid,) + r.f(3))
            materials.append(Material(params, ranges))
        elif r.cstr() != b"NOMAT":
            raise SsfError("expected NOMAT")
        first, last = r.q(2)
        count = last - first
        buffers = {}
        for qid, order, value in quantities:
            if value not in VALUE_COMPONENTS or order not in (0, 1, 2):
                raise SsfError(f"unsupported quantity layout ({qid}, {order}, {value})")
            c = VALUE_COMPONENTS[value]
            dtype = np.int64 if value == INDEX else np.float64
            bufs = []
            for _ in range(order + 1):
                a = r.array(dtype, count * c)
                bufs.append(a.reshape(count, c) if c > 1 else a)
            buffers[qid] = bufs
        blocks.append(Block(first, last, buffers))
    attractors = []
    for _ in range(nattr):
        pos, vel = r.f(4), r.f(4)
        radius, mass = r.f(2)
        attractors.append(Attractor(pos, vel, radius, mass, [_read_param(r) for _ in range(r.q())]))
    if r.o != len(r.d):
        raise SsfError(f"{len(r.d) - r.o} trailing bytes")
    f = SsfFile(run_time, time_step, version, run_type, build_date, wallclock, quantities, materials, blocks,
                attractors)
    if f.particle_count != n:
        raise SsfError(f"the material blocks hold {f.particle_count} particles, the header {n}")
    return f


def write(path, f: SsfFile):
    """BinaryOutput::dump of a file in the layout of its version (attractors from 20210808 on)."""
    out: List[bytes] = []
    nattr = len(f.attractors) if f.version >= 20210808 else 0
    out.append(b"SPH\0")
    out.append(struct.pack("<d3qdq", f.run_time, f.particle_count, len(f.quantities), len(f.materials),
                           f.time_step, f.version))
    for field in (f.run_type, f.build_date):
        if len(field) != 16:
            raise SsfError("run type and build date are 16-byte fields")
        out.append(bytes(field))
    out.append(struct.pack("<2q", f.wallclock, nattr))
    out.append(bytes(PADDING))
    for q in f.quantities:
        out.append(struct.pack("<3q", *q))
    if len(f.blocks) != max(len(f.materials), 1):
        raise SsfError("one particle block per material (or one without materials)")
    for mi, blk in enumerate(f.blocks):
        if f.materials:
            m = f.materials[mi]
            out.append(b"MAT\0" + struct.pack("<2q", mi, len(m.params)))
            for p in m.params:
                _write_param(out, p)
            for rng in m.ranges:
                out.append(struct.pack("<q3d", *rng))
        else:
            out.append(b"NOMAT\0")
        out.append(struct.pack("<2q", blk.first, blk.last))
        for qid, order, value in f.quantities:
            dtype = np.int64 if value == INDEX else np.float64
            for b in blk.buffers[qid][:order + 1]:
                out.append(np.ascontiguousarray(b, dtype=dtype).tobytes())
    for a in f.attractors[:nattr]:
        out.append(struct.pack("<10d", *a.position, *a.velocity, a.radius, a.mass))
        out.append(struct.pack("<q", len(a.params)))
        for p in a.params:
            _write_param(out, p)
    with open(path, "wb") as fh:
        fh.write(b"".join(out))


def _param_dict(m: Material) -> Dict[int, Param]:
    return {p.id: p for p in m.params}


def material_dict(m: Material) -> dict:
    """This package's material parameters from a file's BodySettings (defaults: basalt)."""
    out = basalt()
    ps = _param_dict(m)
    for pid, key in FLOAT_PARAMS.items():
        if pid in ps and ps[pid].type == T_FLOAT:
            out[key] = float(ps[pid].value[0])
    for pid, (lo, hi) in RANGE_PARAMS.items():
        if pid in ps and ps[pid].type == T_INTERVAL:
            out[lo], out[hi] = (float(x) for x in ps[pid].value)
    if B_YIELDING in ps:
        y = int(ps[B_YIELDING].value[0])
        out["yielding"] = {YIELDING_VON_MISES: 1, YIELDING_DRUCKER_PRAGER: 2}.get(y, 0)
    if B_DAMAGE_MODEL in ps:
        out["damage"] = 1 if int(ps[B_DAMAGE_MODEL].value[0]) == FRACTURE_SCALAR_GRADY_KIPP else 0
    return out


# ParticleState field(s) of each quantity of the path, by order
STATE_FIELDS = {POSITION: ("r", "v", "dv"), MASS: ("m",), PRESSURE: ("p",), DENSITY: ("rho", "drho"),
                ENERGY: ("u", "du"), SOUND_SPEED: ("cs",), DEVIATORIC_STRESS: ("s", "ds"),
                DAMAGE: ("damage", "ddamage"), EPS_MIN: ("eps_min",), M_ZERO: ("m_zero",),
                EXPLICIT_GROWTH: ("growth",), N_FLAWS: ("n_flaws",), STRESS_REDUCING: ("reduce",),
                VELOCITY_GRADIENT: ("gradv",), VELOCITY_DIVERGENCE: ("divv",), NEIGHBOR_CNT: ("ncnt",),
                FLAG: ("flag",)}
STATE_VALUES = {POSITION: VECTOR, DEVIATORIC_STRESS: TRACELESS_TENSOR, VELOCITY_GRADIENT: SYMMETRIC_TENSOR,
                N_FLAWS: INDEX, NEIGHBOR_CNT: INDEX, FLAG: INDEX}


def to_state(f: SsfFile) -> ParticleState:
    """The quantities of the path, in storage order; MATERIAL_ID from the material blocks."""
    mats = [material_dict(m) for m in f.materials] or [basalt()]
    st = ParticleState(f.particle_count, mats)
    for mi, blk in enumerate(f.blocks):
        sl = slice(blk.first, blk.last)
        st["mat"][sl] = mi
        for qid, order, value in f.quantities:
            names = STATE_FIELDS.get(qid)
            if names is None:
                continue
            for k, name in enumerate(names[:order + 1]):
                st.a[name][sl] = blk.buffers[qid][k]
    return st


def _float_param(pid, x):
    return Param(pid, T_FLOAT, (float(x),))


def material_params(mat: dict) -> List[Param]:
    """BodySettings entries for one of this package's materials (Tillotson EoS)."""
    ps = [Param(B_EOS, T_ENUM, (EOS_TILLOTSON, 0))]
    defaults = basalt()
    ps += [_float_param(pid, mat.get(key, defaults[key])) for pid, key in sorted(FLOAT_PARAMS.items())]
    ps += [Param(pid, T_INTERVAL, (float(mat[lo]), float(mat[hi]))) for pid, (lo, hi) in sorted(RANGE_PARAMS.items())]
    yielding = {1: YIELDING_VON_MISES, 2: YIELDING_DRUCKER_PRAGER}.get(int(mat.get("yielding", 0)), YIELDING_NONE)
    ps.append(Param(B_YIELDING, T_ENUM, (yielding, 0)))
    ps.append(Param(B_DAMAGE_MODEL, T_ENUM,
                    (FRACTURE_SCALAR_GRADY_KIPP if mat.get("damage") else FRACTURE_NONE, 0)))
    return sorted(ps, key=lambda p: p.id)


def _ranges(mat: dict, qids) -> List[Tuple[int, float, float, float]]:
    known = {DENSITY: (mat["rho_lo"], mat["rho_hi"], mat["rho_min"]), ENERGY: (mat["u_lo"], mat["u_hi"], mat["u_min"]),
             DAMAGE: (mat["d_lo"], mat["d_hi"], mat["d_min"]), DEVIATORIC_STRESS: (-INFTY, INFTY, mat["s_min"])}
    return [(q,) + tuple(float(x) for x in known.get(q, (-INFTY, INFTY, 0.0))) for q in qids]


def from_state(st: ParticleState, run_time: float = 0.0, time_step: float = 0.1, solid: Optional[bool] = None,
               damage: Optional[bool] = None, run_type: str = "SPH", build_date: str = "",
               wallclock: int = 0) -> SsfFile:
    """A file of the state (its particles must be sorted by material, as the reference's storage keeps
    them).  Deviatoric stress and stress reduction are written when a material yields (or `solid`),
    damage quantities when a material fractures (or `damage`)."""
    mat = np.asarray(st["mat"], dtype=np.int64)
    if np.any(np.diff(mat) < 0):
        raise SsfError("particles must be sorted by material")
    solid = any(m.get("yielding") for m in st.materials) if solid is None else solid
    damage = any(m.get("damage") for m in st.materials) if damage is None else damage
    qids = [POSITION, MASS, PRESSURE, DENSITY, ENERGY, SOUND_SPEED]
    if solid:
        qids += [DEVIATORIC_STRESS, STRESS_REDUCING]
    if damage:
        qids += [DAMAGE, EPS_MIN, M_ZERO, EXPLICIT_GROWTH, N_FLAWS]
    qids += [VELOCITY_GRADIENT, VELOCITY_DIVERGENCE, NEIGHBOR_CNT, FLAG]
    orders = {POSITION: 2, DENSITY: 1, ENERGY: 1, DEVIATORIC_STRESS: 1, DAMAGE: 1}
    quantities = [(q, orders.get(q, 0), STATE_VALUES.get(q, SCALAR)) for q in sorted(qids)]
    materials, blocks = [], []
    for mi, m in enumerate(st.materials):
        idx = np.flatnonzero(mat == mi)
        first = int(idx[0]) if idx.size else int(np.searchsorted(mat, mi))
        last = first + idx.size
        buffers = {}
        for qid, order, value in quantities:
            bufs = []
            for name in STATE_FIELDS[qid][:order + 1]:
                a = st.a[name][first:last]
                bufs.append(a.astype(np.int64) if value == INDEX else np.asarray(a, dtype=np.float64))
            buffers[qid] = bufs
        materials.append(Material(material_params(m), _ranges(m, [q for q, _, _ in quantities])))
        blocks.append(Block(first, last, buffers))

    def field(s: str) -> bytes:
        b = s.encode("ascii")
        if len(b) >= 16:
            raise SsfError("run type and build date hold at most 15 characters")
        return b + bytes(16 - len(b))

    return SsfFile(float(run_time), float(time_step), LATEST, field(run_type), field(build_date), int(wallclock),
                   quantities, materials, blocks, [])

"""Binary state files (.ssf): the reference's own files of every format version (core/io/test/Output.cpp
"BinaryInput backward compatibility", written by generateLatestOutput, :316-352) read with the values
that test fixes, rewritten byte for byte, and the package's state written and read back."""
import os

import numpy as np
import pytest

from opensph_amd import initial, ssf
from opensph_amd.state import INFTY

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ssf")
VERSIONS = [0, 20180407, 20181024, 20210320, 20210808]


def path(version):
    return os.path.join(GOLDEN, f"{version}.ssf")


@pytest.mark.parametrize("version", VERSIONS)
def test_reference_files_read_with_their_known_values(version):
    f = ssf.read(path(version))
    assert f.version == version
    assert f.run_time == 20.0 and f.time_step == 1.5
    assert len(f.materials) == 2 and f.particle_count == 311
    # POSITION (2nd order Vector), MASS, PRESSURE, DENSITY (1st), ENERGY (1st), SOUND_SPEED,
    # DEVIATORIC_STRESS (1st, traceless), STRESS_REDUCING, FLAG (index); MATERIAL_ID is not stored
    assert f.quantities == [(0, 2, 1), (1, 0, 0), (2, 0, 0), (3, 1, 0), (4, 1, 0), (5, 0, 0), (6, 1, 4),
                            (16, 0, 0), (42, 0, 5)]
    want = [(1000.0, 0, (1.0, 2.0, 3.0, 0.0), 1), (2000.0, 1, (0.0, 1.0, 2.0, 0.0), 0)]
    for m, blk, (rho0, yielding, center, sph5) in zip(f.materials, f.blocks, want):
        ps = {p.id: p for p in m.params}
        assert ps[ssf.B_DENSITY].value == (rho0,)
        assert ps[ssf.B_YIELDING].value[0] == yielding  # YieldingEnum NONE / ELASTIC
        assert ps[61].value == center                    # BODY_CENTER
        assert ps[4].value == (sph5,)                    # DISTRIBUTE_MODE_SPH5
        assert np.all(blk.buffers[ssf.DENSITY][0] == rho0)
        assert [r[0] for r in m.ranges] == [q[0] for q in f.quantities]
    assert [b.first for b in f.blocks] == [0, f.blocks[0].last] and f.blocks[1].last == 311
    if version >= 20181024:
        assert f.run_type_name() == "sph"
    if version >= 20210808:
        a1, a2 = f.attractors
        assert a1.position == (1.0,) * 4 and a1.velocity == (-1.0,) * 4 and (a1.radius, a1.mass) == (0.5, 2.0)
        assert a2.position == (0.0,) * 4 and a2.velocity == (1.0,) * 4 and (a2.radius, a2.mass) == (0.75, 5.0)
        assert [(p.id, p.value) for p in a1.params] == [(1, (1,))]  # BLACK_HOLE true
        assert [(p.id, p.value) for p in a2.params] == [(1, (0,))]
    else:
        assert f.attractors == []


@pytest.mark.parametrize("version", VERSIONS)
def test_reference_files_rewritten_byte_for_byte(version, tmp_path):
    out = tmp_path / "out.ssf"
    ssf.write(out, ssf.read(path(version)))
    assert out.read_bytes() == open(path(version), "rb").read()


def test_every_version_holds_the_same_buffers():
    """The reference's test compares each version's buffers with the latest output (almostEqual)."""
    files = [ssf.read(path(v)) for v in VERSIONS]
    ref = files[-1]
    for f in files[:-1]:
        for b, rb in zip(f.blocks, ref.blocks):
            for qid in rb.buffers:
                for x, y in zip(b.buffers[qid], rb.buffers[qid]):
                    assert np.allclose(x, y, rtol=1e-12, atol=0.0), qid


def test_state_from_reference_file():
    st = ssf.to_state(ssf.read(path(20210808)))
    assert st.n == 311
    assert np.all(st["mat"][:280] == 0) and np.all(st["mat"][280:] == 1)
    assert np.all(st["rho"][:280] == 1000.0) and np.all(st["rho"][280:] == 2000.0)
    assert st.materials[0]["rho0"] == 1000.0 and st.materials[1]["rho0"] == 2000.0
    assert np.all(st["r"][:, 3] > 0.0)  # smoothing lengths in the h lane


def test_state_written_and_read_back(tmp_path):
    st, cfg = initial.impact(n_target=3000, n_projectile=300, seed=3)
    st.materials = [dict(m, yielding=2, cohesion=8.5e7) for m in st.materials]  # Drucker-Prager
    rng = np.random.default_rng(4)
    for k in ("v", "dv", "drho", "du", "s", "ds", "ddamage", "p", "cs", "gradv", "divv", "reduce"):
        st.a[k][...] = rng.normal(size=st.a[k].shape)
    st["ncnt"][:] = rng.integers(0, 100, st.n)
    f = ssf.from_state(st, run_time=12.5, time_step=0.25, build_date="Dec  2 2026", wallclock=7)
    out = tmp_path / "state.ssf"
    ssf.write(out, f)
    g = ssf.read(out)
    assert (g.run_time, g.time_step, g.version, g.wallclock) == (12.5, 0.25, ssf.LATEST, 7)
    assert g.run_type_name() == "SPH"
    back = ssf.to_state(g)
    for k in ("r", "v", "dv", "m", "rho", "drho", "u", "du", "s", "ds", "damage", "ddamage", "p", "cs",
              "reduce", "gradv", "divv", "eps_min", "m_zero", "growth", "n_flaws", "ncnt", "flag", "mat"):
        assert np.array_equal(back[k], st[k]), k
    for a, b in zip(back.materials, st.materials):
        for key in ("rho0", "A", "B", "a", "b", "alpha", "beta", "u0", "uiv", "ucv", "shear_modulus",
                    "elasticity_limit", "melt_energy", "young_modulus", "rho_lo", "rho_hi", "u_lo", "u_hi",
                    "d_lo", "d_hi", "rho_min", "u_min", "s_min", "d_min", "yielding", "damage", "cohesion",
                    "internal_friction", "dry_friction"):
            assert a[key] == b[key], key
    # the ranges the reference's material reads back (Interval(lower, upper) where lower < upper)
    dens = [r for r in g.materials[0].ranges if r[0] == ssf.DENSITY][0]
    assert dens == (ssf.DENSITY, st.materials[0]["rho_lo"], INFTY, st.materials[0]["rho_min"])
    ssf.write(tmp_path / "again.ssf", g)
    assert (tmp_path / "again.ssf").read_bytes() == out.read_bytes()


def test_malformed_files_raise(tmp_path):
    data = open(path(20210808), "rb").read()
    (tmp_path / "short.ssf").write_bytes(data[:-8])
    with pytest.raises(ssf.SsfError):
        ssf.read(tmp_path / "short.ssf")
    (tmp_path / "magic.ssf").write_bytes(b"SPX" + data[3:])
    with pytest.raises(ssf.SsfError):
        ssf.read(tmp_path / "magic.ssf")
    (tmp_path / "long.ssf").write_bytes(data + b"\0" * 8)
    with pytest.raises(ssf.SsfError):
        ssf.read(tmp_path / "long.ssf")
    st, _ = initial.impact(n_target=500, n_projectile=100, seed=3)
    st["mat"][0] = 1
    with pytest.raises(ssf.SsfError):
        ssf.from_state(st)


def test_file_without_materials_round_trips(tmp_path):
    """The NOMAT layout (a storage without materials: one particle block, no parameters or ranges)."""
    n = 17
    rng = np.random.default_rng(1)
    quantities = [(ssf.POSITION, 2, ssf.VECTOR), (ssf.MASS, 0, ssf.SCALAR), (ssf.FLAG, 0, ssf.INDEX)]
    buffers = {ssf.POSITION: [rng.normal(size=(n, 4)) for _ in range(3)], ssf.MASS: [rng.random(n)],
               ssf.FLAG: [rng.integers(0, 5, n)]}
    f = ssf.SsfFile(3.5, 0.125, ssf.LATEST, b"nbody" + bytes(11), bytes(16), 0, quantities, [],
                    [ssf.Block(0, n, buffers)], [])
    out = tmp_path / "nomat.ssf"
    ssf.write(out, f)
    g = ssf.read(out)
    assert g.materials == [] and g.particle_count == n and g.run_type_name() == "nbody"
    for qid, bufs in buffers.items():
        for a, b in zip(bufs, g.blocks[0].buffers[qid]):
            assert np.array_equal(a, b)
    ssf.write(tmp_path / "again.ssf", g)
    assert (tmp_path / "again.ssf").read_bytes() == out.read_bytes()
    st = ssf.to_state(g)
    assert st.n == n and np.array_equal(st["r"], buffers[ssf.POSITION][0])

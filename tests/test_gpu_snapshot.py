"""Snapshots of GPU-resident fluid (phieng_world3_save / load; snapshot.rs
fluid block :582-636, save :3344-3380, load :4545-4615): the reference's
3D fluid round-trip test on the GPU path, and larger states (several
phases, masses changed by adaptive resolution, more than MAX_PARTICLES_3D
particles)."""
import json

import numpy as np
import pytest

from phi_eng_amd import ffi, scenes
from phi_eng_amd.ffi import FluidConfig3Ffi
from phi_eng_amd.world import World3

pytestmark = pytest.mark.gpu


def _same(a, b):
    pa, va = a.get_state()
    pb, vb = b.get_state()
    ma, ha = a.get_mass()
    mb, hb = b.get_mass()
    return (pa.shape == pb.shape and np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
            and np.array_equal(va.view(np.uint32), vb.view(np.uint32))
            and np.array_equal(ma.view(np.uint32), mb.view(np.uint32)) and np.array_equal(ha, hb))


@pytest.mark.parametrize("binary", [False, True])
def test_reference_fluid_roundtrip_3d(tmp_path, binary):
    """snapshot.rs:8851-8907: a static box and an 18-particle lattice, 60
    steps; save and load give the state bit for bit, and 60 more steps of
    the loaded world equal 60 more steps of the original."""
    w = World3()
    assert w.add_static_box((0.0, -1.0, 0.0), (10.0, 0.5, 10.0)) == 0
    pts = np.array([(ix * 0.25, iy * 0.25, iz * 0.25) for iz in range(2) for iy in range(3) for ix in range(3)],
                   np.float32)
    assert w.add_particles(pts) == 18
    for _ in range(60):
        w.step(1.0 / 60.0)
    w.sync()
    path = str(tmp_path / ("w.bin" if binary else "w.json"))
    assert w.save(path, binary=binary) == 1, ffi.last_error()
    r = World3.load(path)
    assert r is not None, ffi.last_error()
    assert r.count() == 18 and _same(w, r)
    for _ in range(60):
        w.step(1.0 / 60.0)
        r.step(1.0 / 60.0)
    w.sync()
    r.sync()
    assert _same(w, r)
    w.close()
    r.close()


def test_phases_adaptive_masses_and_large_counts(tmp_path):
    """Two phases and adaptive resolution on 12,000 particles (above the
    reference's 4,096 cap; the capacity is lifted on load): masses and
    phases survive both formats and the continuation stays identical."""
    rng = np.random.default_rng(3)
    cfg = FluidConfig3Ffi.default()
    cfg.adaptive_resolution = 1
    cfg.phase_count = 2
    cfg.phase_rest_density[1] = 1.2
    cfg.phase_mass[0] = 0.03
    cfg.phase_mass[1] = 0.04
    w = World3(capacity=40000)
    assert w.set_fluid_config(cfg)
    assert w.add_static_box((2.0, -1.0, 2.0), (6.0, 0.5, 6.0)) == 0
    pos = rng.uniform((0.0, 0.0, 0.0), (4.0, 1.5, 4.0), (12000, 3)).astype(np.float32)
    ph = (np.arange(12000) % 2).astype(np.uint8)
    assert w.add_particles(pos, ph) == 12000
    for _ in range(3):
        w.step(1.0 / 60.0)
    w.sync()
    m, _ = w.get_mass()
    assert len(np.unique(m)) > 2  # merges and splits changed masses
    for binary in (False, True):
        path = str(tmp_path / ("a.bin" if binary else "a.json"))
        assert w.save(path, binary=binary) == 1, ffi.last_error()
        _decoded_equals_state(w, path, binary)
        r = World3.load(path)
        assert r is not None, ffi.last_error()
        assert bytes(r.get_fluid_config()) == bytes(w.get_fluid_config())
        assert _same(w, r)
        r.close()
    r = World3.load(str(tmp_path / "a.bin"))
    for _ in range(3):
        w.step(1.0 / 60.0)
        r.step(1.0 / 60.0)
    w.sync()
    r.sync()
    assert _same(w, r)
    w.close()
    r.close()


def _decoded_equals_state(w, path, binary, dim=3):
    """The file's particle block, read by the independent decoders of
    tests/test_snapshot.py (struct / numpy for PHE1, json for JSON), equals
    the world's state in index order bit for bit: position, velocity, mass
    and phase (FluidParticle3Dto, snapshot.rs:582-592; encode_fluid3
    :6662-6677)."""
    from tests import test_snapshot as ts
    if binary:
        d = ts.decode(path) if dim == 3 else ts.decode2(path)
        got = d["particles"]  # no mass / phase below schema 5 (SNAPSHOT_SCHEMA_AMBIENT)
        count = d["particle_count"]
    else:
        with open(path) as f:
            got = ts.json_particles(json.load(f), dim)
        count = len(got["pos"])
    p, v = w.get_state()
    m, ph = w.get_mass()
    assert count == len(p)
    pairs = [("pos", got["pos"], p), ("vel", got["vel"], v)]
    if "mass" in got:
        pairs.append(("mass", got["mass"], m))
        assert np.array_equal(got["phase"], ph)
    for name, a, b in pairs:
        assert a.shape == b.shape, name
        assert np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32)), name
    return count


@pytest.mark.parametrize("binary", [True, False])
def test_particle_block_of_a_1m_world(tmp_path, binary):
    """A saved 1,048,576-particle dam break: the decoded particle block is
    the GPU state bit for bit.  The count is above MAX_PARTICLES_3D (4,096),
    which the reference's decode_fluid3 refuses (snapshot.rs:6679-6684): such
    a file round-trips on this path only."""
    from tests import test_snapshot as ts
    sc = scenes.c2_dam_break_1m()
    w = scenes.make_world(sc)
    for _ in range(3):
        w.step(sc.dt)
    w.sync()
    path = str(tmp_path / ("m.bin" if binary else "m.json"))
    assert w.save(path, binary=binary) == 1, ffi.last_error()
    count = _decoded_equals_state(w, path, binary)
    assert count > ts.MAX_PARTICLES_3D
    print(f"{count} particles decoded ({'PHE1' if binary else 'JSON'}); beyond the reference's "
          f"{ts.MAX_PARTICLES_3D}-particle limit", flush=True)
    r = World3.load(path)
    assert r is not None, ffi.last_error()
    assert _same(w, r)
    r.close()
    w.close()


@pytest.mark.parametrize("binary", [False, True])
def test_reference_fluid_roundtrip_2d(tmp_path, binary):
    """snapshot.rs:8680-8721 on the GPU path: a 2D box and a 4 x 3 lattice,
    60 steps; the loaded world equals the original and so do 60 more steps
    of each."""
    from phi_eng_amd.world import SHAPE2_BOX, World2
    w = World2()
    assert w.add_body((0.0, -1.0), SHAPE2_BOX, 10.0, 0.5) == 0
    pts = np.array([(ix * 0.25, iy * 0.25) for iy in range(3) for ix in range(4)], np.float32)
    assert w.add_particles(pts) == 12
    for _ in range(60):
        w.step(1.0 / 60.0)
    w.sync()
    path = str(tmp_path / ("w.bin" if binary else "w.json"))
    assert w.save(path, binary=binary) == 1, ffi.last_error()
    _decoded_equals_state(w, path, binary, dim=2)
    r = World2.load(path)
    assert r is not None, ffi.last_error()
    assert r.count() == 12 and _same(w, r)
    for _ in range(60):
        w.step(1.0 / 60.0)
        r.step(1.0 / 60.0)
    w.sync()
    r.sync()
    assert _same(w, r)
    w.close()
    r.close()

"""Snapshots of GPU-resident fluid (phieng_world3_save / load; snapshot.rs
fluid block :582-636, save :3344-3380, load :4545-4615): the reference's
3D fluid round-trip test on the GPU path, and larger states (several
phases, masses changed by adaptive resolution, more than MAX_PARTICLES_3D
particles)."""
import numpy as np
import pytest

from phi_eng_amd import ffi
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

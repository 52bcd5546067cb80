"""C ABI of libphieng_hip.so without a GPU: exports, layouts and the host
logic of the reference's FFI (crates/engine/src/ffi.rs) — validation,
null tolerance, return codes.  No compute call runs here."""
import ctypes
import os
import re

import numpy as np
import pytest

from phi_eng_amd import ffi
from phi_eng_amd.ffi import FluidConfig3Ffi
from phi_eng_amd.world import World3

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            src = open(os.path.join(ROOT, "include", f)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(phieng_\w+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    lib = ffi.lib()
    declared = _declared_functions()
    assert len(declared) >= 20
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(ffi.SIGNATURES), declared ^ set(ffi.SIGNATURES)


def test_fluid_config_layout_matches_ffi_rs():
    """#[repr(C)] FluidConfig3Ffi (ffi.rs:743-767): 112 bytes."""
    assert ctypes.sizeof(FluidConfig3Ffi) == 112
    off = {n: getattr(FluidConfig3Ffi, n).offset for n, _ in FluidConfig3Ffi._fields_}
    assert off["max_substeps"] == 32 and off["adaptive_resolution"] == 36
    assert off["split_density_ratio"] == 40 and off["boundary_particles"] == 48
    assert off["viscosity"] == 52 and off["phase_count"] == 64
    assert off["phase_rest_density"] == 68 and off["phase_mass"] == 84
    assert off["max_velocity"] == 100 and off["particle_radius"] == 104 and off["enabled"] == 108


def test_default_config_round_trip():
    """FluidConfig3::default (fluid3.rs:105-134) and get/set round trip
    (ffi.rs:5841-5900)."""
    with World3() as w:
        c = w.get_fluid_config()
        d = FluidConfig3Ffi.default()
        assert bytes(c) == bytes(d)
        d.stiffness = 123.0
        d.gamma = 7.0
        assert w.set_fluid_config(d)
        assert w.get_fluid_config().stiffness == np.float32(123.0)


@pytest.mark.parametrize("field,value", [
    ("particle_mass", 0.0), ("rest_density", -1.0), ("smoothing_length", 0.0), ("stiffness", -1.0),
    ("gamma", 0.5), ("artificial_viscosity", 2.5), ("surface_tension", -0.1), ("cfl", 1.5),
    ("max_substeps", 0), ("split_density_ratio", 0.0), ("merge_density_ratio", 1.0), ("viscosity", 1.5),
    ("drag", -1.0), ("wall_friction", 1.5), ("phase_count", 0), ("phase_count", 5), ("max_velocity", -1.0),
    ("particle_radius", -0.1), ("stiffness", float("nan")), ("viscosity", float("inf")),
])
def test_invalid_config_rejected_world_unchanged(field, value):
    """fluid3_from_ffi rules (ffi.rs:857-914): 0 returned, config unchanged."""
    with World3() as w:
        before = bytes(w.get_fluid_config())
        c = FluidConfig3Ffi.default()
        setattr(c, field, value)
        assert not w.set_fluid_config(c)
        assert bytes(w.get_fluid_config()) == before


def test_phase_tables_validated():
    with World3() as w:
        c = FluidConfig3Ffi.default()
        c.phase_mass[2] = 0.0
        assert not w.set_fluid_config(c)
        c = FluidConfig3Ffi.default()
        c.phase_rest_density[3] = float("nan")
        assert not w.set_fluid_config(c)


def test_option_flags_read_back_as_bools():
    """adaptive_resolution and boundary_particles are accepted and read back
    as Rust bools; adaptive resolution is refused while a surrogate is
    attached (the surrogate keeps a fixed particle set)."""
    with World3() as w:
        c = FluidConfig3Ffi.default()
        c.adaptive_resolution = 5
        assert w.set_fluid_config(c)
        assert w.get_fluid_config().adaptive_resolution == 1
        c = FluidConfig3Ffi.default()
        c.boundary_particles = 7
        assert w.set_fluid_config(c)
        assert w.get_fluid_config().boundary_particles == 1
        assert w.boundary_info() == (0, 0)  # no ghosts before a step


def test_null_tolerance():
    """Every entry point tolerates a null world (ffi.rs:1-19, 7255-7300)."""
    lib = ffi.lib()
    buf = np.zeros(6, dtype=np.float32)
    fp = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    lib.phieng_world3_free(None)
    lib.phieng_world3_step(None, 0.016)
    lib.phieng_world3_set_gravity(None, 0.0, 0.0, 0.0)
    assert lib.phieng_world3_add_body(None, 0, 0, 0, 1, 1, 1, 1, 1, 1, 0.5, 0) == -1
    assert lib.phieng_world3_set_fluid_config(None, None) == 0
    assert lib.phieng_world3_get_fluid_config(None, None) == 0
    assert lib.phieng_world3_fluid_add(None, fp, 1) == 0
    assert lib.phieng_world3_fluid_add_phase(None, fp, None, 1) == 0
    assert lib.phieng_world3_fluid_count(None) == 0
    assert lib.phieng_world3_fluid_clear(None) == 0
    assert lib.phieng_world3_fluid_get(None, fp, 1) == 0
    assert lib.phieng_world3_fluid_reserve(None, 10) == 0
    assert lib.phieng_world3_fluid_get_mass(None, fp, None, 1) == 0
    assert lib.phieng_world3_sync(None) == 0
    assert lib.phieng_world3_body_count(None) == -1
    assert lib.phieng_world3_get_body_transform(None, 0, fp, fp) == 0
    assert lib.phieng_world3_get_body_velocity(None, 0, fp) == 0
    assert lib.phieng_world3_get_body_angular_velocity(None, 0, fp) == 0
    with World3() as w:
        assert lib.phieng_world3_fluid_add(w.ptr, None, 1) == 0
        assert lib.phieng_world3_fluid_get(w.ptr, None, 1) == 0
        assert lib.phieng_world3_fluid_get_mass(w.ptr, None, None, 1) == 0
        assert lib.phieng_world3_fluid_get_mass(w.ptr, fp, None, 1) == 0  # empty fluid
        assert lib.phieng_world3_get_fluid_config(w.ptr, None) == 0
        lib.phieng_world3_step(w.ptr, float("nan"))
        lib.phieng_world3_step(w.ptr, -1.0)
        lib.phieng_world3_step(w.ptr, 0.016)  # empty fluid: no-op, no device needed
        # neighbour search: 0 cells, 1 rows, 2 automatic; skin in [0, 0.5]
        for v in (0, 1, 2):
            assert lib.phieng_world3_set_neighbour_list(w.ptr, v, 0.1) == 1
        assert lib.phieng_world3_set_neighbour_list(w.ptr, 3, 0.1) == 0
        assert lib.phieng_world3_set_neighbour_list(w.ptr, -1, 0.1) == 0
        assert lib.phieng_world3_set_neighbour_list(w.ptr, 1, 0.6) == 0
        # the surrogate judge's acceleration limit: > 0 (inf turns it off)
        assert lib.phieng_world3_set_surrogate_accel_limit(None, 0.5) == 0
        for bad in (0.0, -1.0, float("nan")):
            assert lib.phieng_world3_set_surrogate_accel_limit(w.ptr, bad) == 0
        assert lib.phieng_world3_set_surrogate_accel_limit(w.ptr, float("inf")) == 1
        assert lib.phieng_world3_surrogate_judge(None, fp, fp, None) == 0
        assert w.surrogate_judge() == (-1.0, -1.0, 0)


def test_add_body_rules():
    """ffi.rs:3703-3760: shape validation, -1 on bad input, ids count up."""
    with World3() as w:
        assert w.add_body((0, 0, 0), 1, 1, 1, 1) == 0
        assert w.add_body((0, 0, 0), 0, 0.5) == 1
        assert w.add_body((0, 0, 0), 2, 0.2, 0.5) == 2
        assert w.add_body((0, 0, 0), 1, 1, -1, 1) == -1
        assert w.add_body((0, 0, 0), 0, 0.0) == -1
        assert w.add_body((0, 0, 0), 7, 1.0) == -1
        assert w.add_body((float("nan"), 0, 0), 0, 1.0) == -1
        assert w.add_body((0, 0, 0), 0, 1.0, density=0.0) == -1
        # a dynamic body (coupled both ways to the fluid) takes the next id
        assert w.add_body((0, 2, 0), 0, 1.0, is_static=0) == 3
        assert w.body_count() == 4
        p, q = w.body_transform(3)
        assert list(p) == [0, 2, 0] and list(q) == [0, 0, 0, 1]
        assert w.body_transform(4) is None
    with World3() as w:
        cfg = FluidConfig3Ffi.default()
        cfg.adaptive_resolution = 1
        assert w.set_fluid_config(cfg)
        # adaptive resolution's substeps do not couple: refused, not ignored
        assert w.add_body((0, 0, 0), 0, 1.0, is_static=0) == -1
        assert "dynamic" in ffi.last_error()
        cfg.adaptive_resolution = 0
        assert w.set_fluid_config(cfg)
        assert w.add_body((0, 0, 0), 0, 1.0, is_static=0) == 0
        cfg.adaptive_resolution = 1
        assert not w.set_fluid_config(cfg)


def test_add_bodies_bulk():
    """ffi.rs:3786-3860 and 1224-1290: every descriptor validated as
    add_body, -1 for a rejected one without rollback, the count added."""
    from phi_eng_amd.ffi import BodyDesc2D, BodyDesc3D
    lib = ffi.lib()
    with World3() as w:
        descs = [dict(pos_x=1.0, pos_y=2.0, pos_z=3.0, shape_tag=1, a=0.5, b=0.5, c=0.5, density=1.0, is_static=1),
                 dict(pos_x=0.0, pos_y=0.0, pos_z=0.0, shape_tag=1, a=0.5, b=-1.0, c=0.5, density=1.0, is_static=1),
                 dict(pos_x=0.0, pos_y=5.0, pos_z=0.0, shape_tag=0, a=0.3, density=2.0, is_static=0),
                 dict(pos_x=float("inf"), pos_y=0.0, pos_z=0.0, shape_tag=0, a=0.3, density=1.0, is_static=1),
                 dict(pos_x=0.0, pos_y=1.0, pos_z=0.0, shape_tag=2, a=0.2, b=0.4, density=1.0, is_static=1),
                 dict(pos_x=0.0, pos_y=1.0, pos_z=0.0, shape_tag=9, a=0.2, density=1.0, is_static=1)]
        added, ids = w.add_bodies(descs)
        assert added == 3
        assert list(ids) == [0, -1, 1, -1, 2, -1]
        assert w.body_count() == 3
        assert list(w.body_transform(0)[0]) == [1, 2, 3]
        assert list(w.body_transform(1)[0]) == [0, 5, 0]
        arr = (BodyDesc3D * 1)()
        out = np.zeros(1, np.int32)
        op = out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        assert lib.phieng_world3_add_bodies(None, arr, 1, op) == 0
        assert lib.phieng_world3_add_bodies(w.ptr, None, 1, op) == 0
        assert lib.phieng_world3_add_bodies(w.ptr, arr, 1, None) == 0
        assert lib.phieng_world3_add_bodies(w.ptr, arr, 0, op) == 0
    w2 = lib.phieng_world_new()
    try:
        arr = (BodyDesc2D * 3)()
        arr[0].pos_x, arr[0].pos_y, arr[0].shape_tag, arr[0].a, arr[0].density, arr[0].is_static = 1, 2, 0, 0.5, 1, 1
        arr[1].pos_x, arr[1].pos_y, arr[1].shape_tag, arr[1].a, arr[1].b, arr[1].density, arr[1].is_static = \
            0, 0, 1, 1.0, 0.5, 1.0, 1
        arr[2].pos_x, arr[2].shape_tag, arr[2].a, arr[2].density, arr[2].is_static = 0, 0, 0.5, 1.0, 0  # dynamic
        out = np.full(3, -2, np.int32)
        n = lib.phieng_world_add_bodies(w2, arr, 3, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        assert n == 2 and list(out) == [0, 1, -1]
        assert "2D" in ffi.last_error()
    finally:
        lib.phieng_world_free(w2)


def test_dynamic_body_free_fall_matches_oracle(oracle_lib):
    """With an empty fluid the step is host-only: the gravity kick and the
    integration of a dynamic body (gravity.rs:355-364, body.rs:1390-1406)
    equal the oracle's bit for bit, and a spinning capsule's rotation too
    (integrate_rotation, compute_cos_sin)."""
    O = oracle_lib
    with World3() as w:
        b = w.add_body((0.0, 5.0, 0.0), 2, 0.2, 0.5, density=2.0, is_static=0)
        ref = [O.rigid(2, (0.0, 5.0, 0.0), 0.2, 0.5, density=2.0, is_static=False)]
        dt = np.float32(1.0 / 60.0)
        for k in range(40):
            w.step(dt)
            O.world_run_f32(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), w.get_fluid_config(),
                            (0.0, -9.81, 0.0), ref, dt, 1)
        p, q = w.body_transform(b)
        v, om = w.body_velocity(b)
        assert np.array_equal(p.view(np.uint32), np.array(ref[0].pos[:], np.float32).view(np.uint32))
        assert np.array_equal(v.view(np.uint32), np.array(ref[0].vel[:], np.float32).view(np.uint32))
        assert list(q) == [0, 0, 0, 1] and list(om) == [0, 0, 0]
        # free fall with linear damping 0.01 (RigidBody3::new): v_k = (v_k-1 + g dt)(1 - 0.01 dt)
        vy = np.float32(0.0)
        for k in range(40):
            vy = np.float32(np.float32(vy + np.float32(-9.81) * np.float32(1.0 * dt)) *
                            np.float32(1.0 - np.float32(0.01 * dt)))
        assert v[1] == vy


def test_fluid_add_rejects_non_finite_atomically():
    """ffi.rs:3592-3594: any non-finite value -> 0 added, nothing changes
    (checked before any device work)."""
    with World3() as w:
        pts = np.zeros((4, 3), dtype=np.float32)
        pts[2, 1] = np.nan
        assert w.add_particles(pts) == 0
        assert w.count() == 0


def test_capacity_defaults_to_reference_cap():
    with World3() as w:
        assert w.reserve(0) == 4096  # MAX_PARTICLES_3D (fluid3.rs:28)
        assert w.reserve(1 << 20) == 1 << 20
        assert w.reserve(1 << 30) == 1 << 24


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(ffi, "_lib", None)
    monkeypatch.setattr(ffi, "lib_path", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError, match="missing"):
        ffi.lib()

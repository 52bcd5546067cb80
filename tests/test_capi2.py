"""The 2D C ABI (phieng_world_*, include/phieng_fluid2.h) without a GPU:
validation, null tolerance and return codes of the reference's 2D exports
(crates/engine/src/ffi.rs:51-94, 999-1026, 1159-1203, 2486-2594).  No
compute call runs here."""
import ctypes

import numpy as np
import pytest

from phi_eng_amd import ffi
from phi_eng_amd.ffi import FluidConfig3Ffi
from phi_eng_amd.world import SHAPE2_BOX, SHAPE2_CIRCLE, World2


def test_default_config_round_trip():
    """FluidConfig2::default (fluid.rs:142-169) equals the 3D default field
    for field; get/set round trip."""
    with World2() as w:
        c = w.get_fluid_config()
        assert bytes(c) == bytes(FluidConfig3Ffi.default())
        c.stiffness = 57.0
        c.gamma = 7.0
        assert w.set_fluid_config(c)
        assert w.get_fluid_config().gamma == np.float32(7.0)


@pytest.mark.parametrize("field,value", [
    ("particle_mass", 0.0), ("smoothing_length", -1.0), ("gamma", 0.9), ("cfl", 2.0), ("max_substeps", 0),
    ("merge_density_ratio", 0.5), ("wall_friction", -0.1), ("phase_count", 0), ("rest_density", float("nan")),
])
def test_invalid_config_rejected_world_unchanged(field, value):
    """fluid2_from_ffi (ffi.rs:773-840): 0 returned, config unchanged."""
    with World2() as w:
        before = bytes(w.get_fluid_config())
        c = FluidConfig3Ffi.default()
        setattr(c, field, value)
        assert not w.set_fluid_config(c)
        assert bytes(w.get_fluid_config()) == before


def test_option_flags_accepted():
    with World2() as w:
        c = FluidConfig3Ffi.default()
        c.adaptive_resolution = 1
        assert w.set_fluid_config(c)
        assert w.get_fluid_config().adaptive_resolution == 1
        c = FluidConfig3Ffi.default()
        c.boundary_particles = 1
        assert w.set_fluid_config(c)
        assert w.boundary_info() == (0, 0)


def test_add_body_rules():
    """ffi.rs:1159-1203: circle needs a > 0, box a, b > 0, density > 0;
    ids count up; a dynamic body is refused on the GPU path."""
    with World2() as w:
        assert w.add_body((0, 0), SHAPE2_BOX, 1.0, 0.5) == 0
        assert w.add_body((1, 2), SHAPE2_CIRCLE, 0.3) == 1
        assert w.add_body((0, 0), SHAPE2_BOX, 1.0, 0.0) == -1
        assert w.add_body((0, 0), SHAPE2_CIRCLE, -1.0) == -1
        assert w.add_body((0, 0), 2, 1.0, 1.0) == -1  # no capsule tag in 2D
        assert w.add_body((float("inf"), 0), SHAPE2_CIRCLE, 1.0) == -1
        assert w.add_body((0, 0), SHAPE2_CIRCLE, 1.0, density=-2.0) == -1
        assert w.add_body((0, 0), SHAPE2_CIRCLE, 1.0, is_static=0) == -1
        assert "dynamic" in ffi.last_error()
        assert w.add_body((0, 0), SHAPE2_CIRCLE, 1.0) == 2


def test_null_tolerance():
    lib = ffi.lib()
    buf = np.zeros(4, dtype=np.float32)
    fp = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    lib.phieng_world_free(None)
    lib.phieng_world_step(None, 0.016)
    lib.phieng_world_set_gravity(None, 0.0, 1.0)
    assert lib.phieng_world_add_body(None, 0, 0, 0, 1, 0, 1, 1, 0.5, 0) == -1
    assert lib.phieng_world_set_fluid_config(None, None) == 0
    assert lib.phieng_world_get_fluid_config(None, None) == 0
    assert lib.phieng_world_fluid_add(None, fp, 1) == 0
    assert lib.phieng_world_fluid_add_phase(None, fp, None, 1) == 0
    assert lib.phieng_world_fluid_count(None) == 0
    assert lib.phieng_world_fluid_clear(None) == 0
    assert lib.phieng_world_fluid_get(None, fp, 1) == 0
    assert lib.phieng_world_fluid_reserve(None, 10) == 0
    assert lib.phieng_world_fluid_set_velocity(None, fp, 1) == 0
    assert lib.phieng_world_fluid_set_record(None, 1) == 0
    assert lib.phieng_world_fluid_get_record(None, None, None, None, None, None, None, 1) == 0
    assert lib.phieng_world_sync(None) == 0
    with World2() as w:
        assert lib.phieng_world_fluid_add(w.ptr, None, 1) == 0
        assert lib.phieng_world_fluid_get(w.ptr, None, 1) == 0
        assert lib.phieng_world_set_fluid_config(w.ptr, None) == 0
        lib.phieng_world_step(w.ptr, float("nan"))
        lib.phieng_world_step(w.ptr, 0.0)
        lib.phieng_world_step(w.ptr, 0.016)  # empty fluid: no-op, no device needed
        lib.phieng_world_set_gravity(w.ptr, float("nan"), 1.0)  # ignored
        assert w.sync() is None


def test_fluid_add_rejects_non_finite_atomically():
    """ffi.rs:2496-2498: any non-finite value -> 0 added."""
    with World2() as w:
        pts = np.zeros((5, 2), dtype=np.float32)
        pts[3, 0] = np.inf
        assert w.add_particles(pts) == 0
        assert w.count() == 0


def test_capacity_defaults_to_reference_cap():
    with World2() as w:
        assert w.reserve(0) == 4096  # MAX_PARTICLES_2D (fluid.rs:26)
        assert w.reserve(1 << 16) == 1 << 16

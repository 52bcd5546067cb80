"""ctypes binding of ``libphieng_hip.so`` (include/phieng_fluid3.h, phieng_fluid2.h).

Loading fails loudly when the library has not been built: there is no
fallback implementation of the fluid step in Python.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "libphieng_hip.so")


class FluidConfig3Ffi(ctypes.Structure):
    """``#[repr(C)] FluidConfig3Ffi`` (crates/engine/src/ffi.rs:743-767), 112 bytes."""

    _fields_ = [
        ("particle_mass", ctypes.c_float),
        ("rest_density", ctypes.c_float),
        ("smoothing_length", ctypes.c_float),
        ("stiffness", ctypes.c_float),
        ("gamma", ctypes.c_float),
        ("artificial_viscosity", ctypes.c_float),
        ("surface_tension", ctypes.c_float),
        ("cfl", ctypes.c_float),
        ("max_substeps", ctypes.c_uint32),
        ("adaptive_resolution", ctypes.c_uint8),
        ("split_density_ratio", ctypes.c_float),
        ("merge_density_ratio", ctypes.c_float),
        ("boundary_particles", ctypes.c_uint8),
        ("viscosity", ctypes.c_float),
        ("drag", ctypes.c_float),
        ("wall_friction", ctypes.c_float),
        ("phase_count", ctypes.c_uint8),
        ("phase_rest_density", ctypes.c_float * 4),
        ("phase_mass", ctypes.c_float * 4),
        ("max_velocity", ctypes.c_float),
        ("particle_radius", ctypes.c_float),
        ("enabled", ctypes.c_uint8),
    ]

    @classmethod
    def default(cls):
        """FluidConfig3::default (crates/dynamics/src/fluid3.rs:105-134)."""
        c = cls()
        c.particle_mass = 0.0625
        c.rest_density = 1.0
        c.smoothing_length = 0.5
        c.stiffness = 100.0
        c.gamma = 1.0
        c.artificial_viscosity = 0.0
        c.surface_tension = 0.0
        c.cfl = 0.0
        c.max_substeps = 1
        c.adaptive_resolution = 0
        c.split_density_ratio = 0.5
        c.merge_density_ratio = 1.6
        c.boundary_particles = 0
        c.viscosity = 0.3
        c.drag = 4.0
        c.wall_friction = 0.0
        c.phase_count = 1
        for k in range(4):
            c.phase_rest_density[k] = 1.0
            c.phase_mass[k] = 0.0625
        c.max_velocity = 30.0
        c.particle_radius = 0.125
        c.enabled = 1
        return c

    def copy(self):
        c = FluidConfig3Ffi()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(self))
        return c


assert ctypes.sizeof(FluidConfig3Ffi) == 112


class BodyDesc3D(ctypes.Structure):
    """#[repr(C)] BodyDesc3D (crates/engine/src/ffi.rs:3765-3784), 44 bytes."""
    _fields_ = [("pos_x", ctypes.c_float), ("pos_y", ctypes.c_float), ("pos_z", ctypes.c_float),
                ("shape_tag", ctypes.c_uint32), ("a", ctypes.c_float), ("b", ctypes.c_float),
                ("c", ctypes.c_float), ("density", ctypes.c_float), ("is_static", ctypes.c_uint8),
                ("_pad", ctypes.c_uint8 * 3), ("friction", ctypes.c_float), ("restitution", ctypes.c_float)]


class BodyDesc2D(ctypes.Structure):
    """#[repr(C)] BodyDesc2D (crates/engine/src/ffi.rs:1205-1222), 36 bytes."""
    _fields_ = [("pos_x", ctypes.c_float), ("pos_y", ctypes.c_float), ("shape_tag", ctypes.c_uint32),
                ("a", ctypes.c_float), ("b", ctypes.c_float), ("density", ctypes.c_float),
                ("is_static", ctypes.c_uint8), ("_pad", ctypes.c_uint8 * 3), ("friction", ctypes.c_float),
                ("restitution", ctypes.c_float)]


assert ctypes.sizeof(BodyDesc3D) == 44 and ctypes.sizeof(BodyDesc2D) == 36

_P = ctypes.c_void_p
_F = ctypes.c_float
_U32 = ctypes.c_uint32
_I32 = ctypes.c_int32
_fp = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_cfgp = ctypes.POINTER(FluidConfig3Ffi)

# name -> (restype, argtypes); every symbol include/phieng_fluid3.h and
# include/phieng_fluid2.h declare.
SIGNATURES = {
    "phieng_world3_new": (_P, []),
    "phieng_world3_free": (None, [_P]),
    "phieng_world3_step": (None, [_P, _F]),
    "phieng_world3_set_gravity": (None, [_P, _F, _F, _F]),
    "phieng_world3_add_body": (_I32, [_P, _F, _F, _F, _U32, _F, _F, _F, _F, ctypes.c_uint8, _F, _F]),
    "phieng_world3_add_bodies": (_U32, [_P, ctypes.POINTER(BodyDesc3D), _U32, _i32p]),
    "phieng_world3_body_count": (_I32, [_P]),
    "phieng_world3_get_body_transform": (_I32, [_P, _U32, _fp, _fp]),
    "phieng_world3_get_body_velocity": (_I32, [_P, _U32, _fp]),
    "phieng_world3_get_body_angular_velocity": (_I32, [_P, _U32, _fp]),
    "phieng_world3_set_fluid_config": (_I32, [_P, _cfgp]),
    "phieng_world3_get_fluid_config": (_I32, [_P, _cfgp]),
    "phieng_world3_fluid_add": (_U32, [_P, _fp, _U32]),
    "phieng_world3_fluid_add_phase": (_U32, [_P, _fp, _u8p, _U32]),
    "phieng_world3_fluid_count": (_U32, [_P]),
    "phieng_world3_fluid_clear": (_I32, [_P]),
    "phieng_world3_fluid_get": (_U32, [_P, _fp, _U32]),
    "phieng_world3_save": (_I32, [_P, ctypes.c_char_p, ctypes.c_size_t, _U32, _U32]),
    "phieng_world3_load": (_P, [ctypes.c_char_p, ctypes.c_size_t, _U32]),
    "phieng_world3_fluid_reserve": (_U32, [_P, _U32]),
    "phieng_world3_fluid_get_mass": (_U32, [_P, _fp, _u8p, _U32]),
    "phieng_world3_fluid_set_velocity": (_U32, [_P, _fp, _U32]),
    "phieng_world3_fluid_set_record": (_I32, [_P, _U32]),
    "phieng_world3_fluid_get_record": (_U32, [_P, _i32p, _u32p, _fp, _fp, _fp, _u32p, _U32]),
    "phieng_world3_set_timing": (_I32, [_P, _I32]),
    "phieng_world3_get_timing": (_I32, [_P, _dp, _u64p, _I32]),
    "phieng_kernel_name": (ctypes.c_char_p, [_I32]),
    "phieng_world3_sync": (_I32, [_P]),
    "phieng_last_error": (ctypes.c_char_p, []),
    "phieng_batch3_new": (_P, [ctypes.POINTER(_P), _U32]),
    "phieng_batch3_free": (None, [_P]),
    "phieng_batch3_set_mode": (_I32, [_P, _I32, _U32]),
    "phieng_batch3_set_spin_limit": (_I32, [_P, _U32]),
    "phieng_batch3_step": (_I32, [_P, _F]),
    "phieng_batch3_sync": (_I32, [_P]),
    "phieng_batch3_sync_worlds": (_I32, [_P]),
    "phieng_batch3_count": (_U32, [_P]),
    "phieng_batch3_autotune": (_I32, [_P, _F, _U32, ctypes.c_char_p, _u32p, _dp]),
    "phieng_batch3_set_timing": (_I32, [_P, _I32]),
    "phieng_batch3_get_timing": (_I32, [_P, _dp, _u64p, _I32]),
    "phieng_batch3_last_error": (ctypes.c_char_p, []),
    "phieng_surrogate3_new": (_P, [_U32]),
    "phieng_surrogate3_free": (None, [_P]),
    "phieng_surrogate3_num_params": (_U32, []),
    "phieng_surrogate3_get_params": (_I32, [_P, _fp, _U32]),
    "phieng_surrogate3_set_params": (_I32, [_P, _fp, _U32]),
    "phieng_surrogate3_mlp_forward": (_I32, [_P, _fp, _U32, _fp]),
    "phieng_surrogate3_mlp_forward_noalloc": (_I32, [_P, _fp, _U32, _fp]),
    "phieng_surrogate3_mlp_backward": (_I32, [_P, _fp, _fp, _U32, _fp]),
    "phieng_surrogate3_add_samples": (_I32, [_P, _fp, _U32]),
    "phieng_surrogate3_num_samples": (_U32, [_P]),
    "phieng_surrogate3_clear_samples": (_I32, [_P]),
    "phieng_surrogate3_get_samples": (_I32, [_P, _fp, _U32, _U32]),
    "phieng_surrogate3_train": (_I32, [_P, _U32, _U32, _F, _U32, _fp, _fp, _fp]),
    "phieng_surrogate3_eval": (_I32, [_P, _fp, _fp]),
    "phieng_surrogate3_last_error": (ctypes.c_char_p, []),
    "phieng_world3_set_surrogate": (_I32, [_P, _P, _U32, _F, _I32]),
    "phieng_world3_collect_samples": (_I32, [_P, _P, _U32, _U32]),
    "phieng_world3_surrogate_stats": (_I32, [_P, _u64p, _u64p, _u64p, _fp]),
    "phieng_world3_set_surrogate_skin": (_I32, [_P, _F]),
    "phieng_world3_set_surrogate_accel_limit": (_I32, [_P, _F]),
    "phieng_world3_surrogate_judge": (_I32, [_P, _fp, _fp, _u64p]),
    "phieng_world3_boundary_info": (_I32, [_P, _u32p, _u64p]),
    "phieng_world3_set_neighbour_list": (_I32, [_P, _I32, _F]),
    "phieng_world3_neighbour_stats": (_I32, [_P, _u64p, _u64p, _u64p, _u64p]),
    "phieng_world3_boundary_get": (_U32, [_P, _fp, _fp, _fp, _U32]),
    # include/phieng_fluid2.h (World2; FluidConfig2Ffi has FluidConfig3Ffi's layout)
    "phieng_world_new": (_P, []),
    "phieng_world_free": (None, [_P]),
    "phieng_world_step": (None, [_P, _F]),
    "phieng_world_set_gravity": (None, [_P, _F, _F]),
    "phieng_world_add_body": (_I32, [_P, _F, _F, _U32, _F, _F, _F, ctypes.c_uint8, _F, _F]),
    "phieng_world_add_bodies": (_U32, [_P, ctypes.POINTER(BodyDesc2D), _U32, _i32p]),
    "phieng_world_set_fluid_config": (_I32, [_P, _cfgp]),
    "phieng_world_get_fluid_config": (_I32, [_P, _cfgp]),
    "phieng_world_fluid_add": (_U32, [_P, _fp, _U32]),
    "phieng_world_fluid_add_phase": (_U32, [_P, _fp, _u8p, _U32]),
    "phieng_world_fluid_count": (_U32, [_P]),
    "phieng_world_fluid_clear": (_I32, [_P]),
    "phieng_world_fluid_get": (_U32, [_P, _fp, _U32]),
    "phieng_world_save": (_I32, [_P, ctypes.c_char_p, ctypes.c_size_t, _U32, _U32]),
    "phieng_world_load": (_P, [ctypes.c_char_p, ctypes.c_size_t, _U32]),
    "phieng_world_fluid_reserve": (_U32, [_P, _U32]),
    "phieng_world_fluid_get_mass": (_U32, [_P, _fp, _u8p, _U32]),
    "phieng_world_fluid_set_velocity": (_U32, [_P, _fp, _U32]),
    "phieng_world_fluid_set_record": (_I32, [_P, _U32]),
    "phieng_world_fluid_get_record": (_U32, [_P, _i32p, _u32p, _fp, _fp, _fp, _u32p, _U32]),
    "phieng_world_sync": (_I32, [_P]),
    "phieng_world_boundary_info": (_I32, [_P, _u32p, _u64p]),
    "phieng_world_set_neighbour_list": (_I32, [_P, _I32, _F]),
    "phieng_world_neighbour_stats": (_I32, [_P, _u64p, _u64p, _u64p, _u64p]),
    "phieng_world_boundary_get": (_U32, [_P, _fp, _fp, _fp, _U32]),
}

_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(lib_path):
            raise RuntimeError(
                f"{lib_path} is missing: build it with `python -m phi_eng_amd.build` "
                "(the fluid step has no non-HIP implementation)")
        l = ctypes.CDLL(lib_path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def last_error():
    return lib().phieng_last_error().decode()

"""Python mirror of the reference World3 / World2 fluid surfaces over the C ABI.

Method names follow the reference (``World3::step``, ``fluid.add_particles``,
``phieng_world3_*`` in crates/engine/src/ffi.rs); arrays are numpy.
"""
import ctypes
import os

import numpy as np

from .ffi import FluidConfig3Ffi, last_error, lib

SHAPE3_SPHERE, SHAPE3_BOX, SHAPE3_CAPSULE = 0, 1, 2
RECORD_CELLS, RECORD_DENSITY, RECORD_ACCEL = 1, 2, 4


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _i32ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


class World3:
    """Owns one ``World3*`` from ``phieng_world3_new`` (ffi.rs:3199)."""

    def __init__(self, capacity=None, _ptr=None):
        self._l = lib()
        self.ptr = _ptr if _ptr else self._l.phieng_world3_new()
        if not self.ptr:
            raise MemoryError("phieng_world3_new returned null")
        if capacity is not None:
            self.reserve(capacity)

    # -- snapshots (ffi.rs:4036-4117) -----------------------------------
    def save(self, path, binary=False, fresh=False):
        """phieng_world3_save; returns its code (1, 0, -1 or -2)."""
        b = os.fsencode(path)
        return self._l.phieng_world3_save(self.ptr, b, len(b), int(binary), int(fresh))

    @classmethod
    def load(cls, path, fresh=False):
        """phieng_world3_load; a new World3, or None (last_error says why)."""
        b = os.fsencode(path)
        ptr = lib().phieng_world3_load(b, len(b), int(fresh))
        return cls(_ptr=ptr) if ptr else None

    def close(self):
        if self.ptr:
            self._l.phieng_world3_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- configuration --------------------------------------------------
    def set_gravity(self, g):
        self._l.phieng_world3_set_gravity(self.ptr, float(g[0]), float(g[1]), float(g[2]))

    def set_fluid_config(self, cfg: FluidConfig3Ffi) -> bool:
        return bool(self._l.phieng_world3_set_fluid_config(self.ptr, ctypes.byref(cfg)))

    def get_fluid_config(self) -> FluidConfig3Ffi:
        c = FluidConfig3Ffi()
        if not self._l.phieng_world3_get_fluid_config(self.ptr, ctypes.byref(c)):
            raise RuntimeError("get_fluid_config failed")
        return c

    def add_static_box(self, center, half_extents):
        return self._l.phieng_world3_add_body(self.ptr, *map(float, center), SHAPE3_BOX,
                                              *map(float, half_extents), 1.0, 1, 0.5, 0.0)

    def add_body(self, center, shape, a, b=0.0, c=0.0, density=1.0, is_static=1, friction=0.5, restitution=0.0):
        return self._l.phieng_world3_add_body(self.ptr, *map(float, center), int(shape), float(a),
                                              float(b), float(c), float(density), int(is_static),
                                              float(friction), float(restitution))

    def add_bodies(self, descs):
        """descs: sequence of dicts with the BodyDesc3D fields; returns (added, ids)."""
        from .ffi import BodyDesc3D
        arr = (BodyDesc3D * max(1, len(descs)))()
        for k, d in enumerate(descs):
            for key, v in d.items():
                setattr(arr[k], key, v)
        ids = np.full(max(1, len(descs)), -2, np.int32)
        added = self._l.phieng_world3_add_bodies(self.ptr, arr, len(descs), _i32ptr(ids))
        return added, ids[:len(descs)]

    def body_count(self):
        return self._l.phieng_world3_body_count(self.ptr)

    def body_transform(self, body_id):
        """(position (3,), quaternion x, y, z, w (4,)) or None for an unknown id."""
        p = np.zeros(3, np.float32)
        q = np.zeros(4, np.float32)
        if not self._l.phieng_world3_get_body_transform(self.ptr, int(body_id), _fptr(p), _fptr(q)):
            return None
        return p, q

    def body_velocity(self, body_id):
        v = np.zeros(3, np.float32)
        w = np.zeros(3, np.float32)
        if not self._l.phieng_world3_get_body_velocity(self.ptr, int(body_id), _fptr(v)):
            return None
        self._l.phieng_world3_get_body_angular_velocity(self.ptr, int(body_id), _fptr(w))
        return v, w

    def reserve(self, capacity):
        return self._l.phieng_world3_fluid_reserve(self.ptr, int(capacity))

    # -- particles -------------------------------------------------------
    def add_particles(self, positions, phases=None):
        p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        total = 0
        for s in range(0, len(p), 1_000_000):  # ffi.rs:3587 caps each call
            chunk = np.ascontiguousarray(p[s:s + 1_000_000])
            if phases is None:
                got = self._l.phieng_world3_fluid_add(self.ptr, _fptr(chunk), len(chunk))
            else:
                ph = np.ascontiguousarray(np.asarray(phases, dtype=np.uint8)[s:s + 1_000_000])
                got = self._l.phieng_world3_fluid_add_phase(
                    self.ptr, _fptr(chunk), ph.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                    len(chunk))
            total += got
            if got < len(chunk):
                break
        return total

    def count(self):
        return self._l.phieng_world3_fluid_count(self.ptr)

    def clear(self):
        return self._l.phieng_world3_fluid_clear(self.ptr)

    def get_state(self):
        """(positions, velocities) in insertion order, each (N, 3) float32."""
        n = self.count()
        out = np.empty((n, 6), dtype=np.float32)
        got = self._l.phieng_world3_fluid_get(self.ptr, _fptr(out), n)
        if got != n:
            raise RuntimeError(f"fluid_get returned {got} of {n}: {last_error()}")
        return out[:, :3].copy(), out[:, 3:].copy()

    def get_mass(self):
        """(masses float32, phases uint8) in insertion order."""
        n = self.count()
        m = np.empty(n, dtype=np.float32)
        ph = np.empty(n, dtype=np.uint8)
        if n == 0:
            return m, ph
        got = self._l.phieng_world3_fluid_get_mass(self.ptr, _fptr(m), ph.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), n)
        if got != n:
            raise RuntimeError(f"fluid_get_mass returned {got} of {n}: {last_error()}")
        return m, ph

    def set_velocity(self, vel):
        v = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
        return self._l.phieng_world3_fluid_set_velocity(self.ptr, _fptr(v), len(v))

    # -- stepping --------------------------------------------------------
    def step(self, dt):
        self._l.phieng_world3_step(self.ptr, float(dt))

    def sync(self):
        if not self._l.phieng_world3_sync(self.ptr):
            raise RuntimeError(last_error())

    # -- parity readback ---------------------------------------------------
    def set_record(self, flags):
        if not self._l.phieng_world3_fluid_set_record(self.ptr, int(flags)):
            raise RuntimeError(last_error())

    def get_record(self):
        n = self.count()
        cells = np.zeros((n, 3), dtype=np.int32)
        order = np.zeros(n, dtype=np.uint32)
        rho = np.zeros(n, dtype=np.float32)
        p = np.zeros(n, dtype=np.float32)
        acc = np.zeros((n, 3), dtype=np.float32)
        cnt = np.zeros(n, dtype=np.uint32)
        got = self._l.phieng_world3_fluid_get_record(
            self.ptr, cells.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            order.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _fptr(rho), _fptr(p), _fptr(acc),
            cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n)
        if got != n:
            raise RuntimeError(f"get_record returned {got}: {last_error()}")
        return dict(cells=cells, order=order, density=rho, pressure=p, accel=acc, count=cnt)

    # -- boundary ghosts ---------------------------------------------------
    def boundary_info(self):
        c, r = ctypes.c_uint32(), ctypes.c_uint64()
        self._l.phieng_world3_boundary_info(self.ptr, ctypes.byref(c), ctypes.byref(r))
        return c.value, r.value

    def set_neighbour_list(self, verlet=True, skin=0.1):
        """verlet: True (the row form), False (the cells), or "auto" (the
        default: rows from 131,072 particles up)."""
        v = 2 if verlet == "auto" else (1 if verlet else 0)
        if not self._l.phieng_world3_set_neighbour_list(self.ptr, v, float(skin)):
            raise ValueError("neighbour list settings rejected")

    def neighbour_stats(self):
        a, b, c, d = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        if not self._l.phieng_world3_neighbour_stats(self.ptr, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d)):
            raise RuntimeError(last_error())
        return {"substeps": a.value, "rebuilds": b.value, "repaired": c.value, "list_substeps": d.value}

    def boundary(self):
        """Ghosts in generation order: dict(pos (G, 3), density, pressure)."""
        g, _ = self.boundary_info()
        xyz = np.zeros((g, 3), np.float32)
        rho = np.zeros(g, np.float32)
        p = np.zeros(g, np.float32)
        got = self._l.phieng_world3_boundary_get(self.ptr, _fptr(xyz), _fptr(rho), _fptr(p), g)
        if got != g:
            raise RuntimeError(f"boundary_get returned {got} of {g}: {last_error()}")
        return dict(pos=xyz, density=rho, pressure=p)

    # -- learned pressure term ---------------------------------------------
    def set_surrogate(self, sur, check_every=10, max_density_error=0.01, use_network=True):
        ok = self._l.phieng_world3_set_surrogate(self.ptr, sur.ptr if sur is not None else None,
                                                 int(check_every), float(max_density_error),
                                                 1 if use_network else 0)
        if not ok:
            raise RuntimeError("set_surrogate rejected its arguments")

    def collect_samples(self, sur, every=10, stride=64):
        if not self._l.phieng_world3_collect_samples(self.ptr, sur.ptr if sur is not None else None,
                                                     int(every), int(stride)):
            raise RuntimeError("collect_samples rejected its arguments")

    def set_surrogate_skin(self, skin):
        if not self._l.phieng_world3_set_surrogate_skin(self.ptr, float(skin)):
            raise RuntimeError("set_surrogate_skin rejected its argument")

    def set_surrogate_accel_limit(self, limit):
        if not self._l.phieng_world3_set_surrogate_accel_limit(self.ptr, float(limit)):
            raise RuntimeError("set_surrogate_accel_limit rejected its argument")

    def surrogate_judge(self):
        """Newest judged (density error, acceleration error), -1 when absent,
        and the count of judges that kept the order and list of the one before."""
        d, a, k = ctypes.c_float(), ctypes.c_float(), ctypes.c_uint64()
        if not self._l.phieng_world3_surrogate_judge(self.ptr, ctypes.byref(d), ctypes.byref(a), ctypes.byref(k)):
            raise RuntimeError(self._l.phieng_last_error().decode())
        return d.value, a.value, k.value

    def surrogate_stats(self):
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        e = ctypes.c_float()
        self._l.phieng_world3_surrogate_stats(self.ptr, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                              ctypes.byref(e))
        return dict(surrogate_steps=a.value, plain_steps=b.value, fallbacks=c.value, last_density_error=e.value)

    # -- timing -----------------------------------------------------------
    def set_timing(self, on):
        self._l.phieng_world3_set_timing(self.ptr, 1 if on else 0)

    def get_timing(self):
        k = 32
        ms = (ctypes.c_double * k)()
        cnt = (ctypes.c_uint64 * k)()
        m = self._l.phieng_world3_get_timing(self.ptr, ms, cnt, k)
        out = {}
        for i in range(min(m, k)):
            if cnt[i]:
                out[self._l.phieng_kernel_name(i).decode()] = (ms[i], cnt[i])
        return out


SHAPE2_CIRCLE, SHAPE2_BOX = 0, 1


class World2:
    """Owns one ``World2*`` from ``phieng_world_new`` (ffi.rs:51-53): the 2D
    fluid surface (``phieng_world_*``), arrays of xy pairs."""

    def __init__(self, capacity=None, _ptr=None):
        self._l = lib()
        self.ptr = _ptr if _ptr else self._l.phieng_world_new()
        if not self.ptr:
            raise MemoryError("phieng_world_new returned null")
        if capacity is not None:
            self.reserve(capacity)

    def save(self, path, binary=False, fresh=False):
        """phieng_world_save; returns its code (1, 0, -1 or -2)."""
        b = os.fsencode(path)
        return self._l.phieng_world_save(self.ptr, b, len(b), int(binary), int(fresh))

    @classmethod
    def load(cls, path, fresh=False):
        """phieng_world_load; a new World2, or None (last_error says why)."""
        b = os.fsencode(path)
        ptr = lib().phieng_world_load(b, len(b), int(fresh))
        return cls(_ptr=ptr) if ptr else None

    def close(self):
        if self.ptr:
            self._l.phieng_world_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_gravity(self, g):
        self._l.phieng_world_set_gravity(self.ptr, float(g[0]), float(g[1]))

    def set_fluid_config(self, cfg: FluidConfig3Ffi) -> bool:
        return bool(self._l.phieng_world_set_fluid_config(self.ptr, ctypes.byref(cfg)))

    def get_fluid_config(self) -> FluidConfig3Ffi:
        c = FluidConfig3Ffi()
        if not self._l.phieng_world_get_fluid_config(self.ptr, ctypes.byref(c)):
            raise RuntimeError("get_fluid_config failed")
        return c

    def add_body(self, center, shape, a, b=0.0, density=1.0, is_static=1):
        return self._l.phieng_world_add_body(self.ptr, float(center[0]), float(center[1]), int(shape), float(a),
                                             float(b), float(density), int(is_static), 0.5, 0.0)

    def add_static_box(self, center, half_extents):
        return self.add_body(center, SHAPE2_BOX, half_extents[0], half_extents[1])

    def reserve(self, capacity):
        return self._l.phieng_world_fluid_reserve(self.ptr, int(capacity))

    def add_particles(self, positions, phases=None):
        p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 2)
        if phases is None:
            return self._l.phieng_world_fluid_add(self.ptr, _fptr(p), len(p))
        ph = np.ascontiguousarray(phases, dtype=np.uint8)
        return self._l.phieng_world_fluid_add_phase(self.ptr, _fptr(p),
                                                    ph.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), len(p))

    def count(self):
        return self._l.phieng_world_fluid_count(self.ptr)

    def clear(self):
        return self._l.phieng_world_fluid_clear(self.ptr)

    def get_state(self):
        """(positions, velocities) in insertion order, each (N, 2) float32."""
        n = self.count()
        out = np.empty((n, 4), dtype=np.float32)
        got = self._l.phieng_world_fluid_get(self.ptr, _fptr(out), n)
        if got != n:
            raise RuntimeError(f"fluid_get returned {got} of {n}: {last_error()}")
        return out[:, :2].copy(), out[:, 2:].copy()

    def get_mass(self):
        """(masses float32, phases uint8) in insertion order."""
        n = self.count()
        m = np.empty(n, dtype=np.float32)
        ph = np.empty(n, dtype=np.uint8)
        if n == 0:
            return m, ph
        got = self._l.phieng_world_fluid_get_mass(self.ptr, _fptr(m), ph.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), n)
        if got != n:
            raise RuntimeError(f"fluid_get_mass returned {got} of {n}: {last_error()}")
        return m, ph

    def set_velocity(self, vel):
        v = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 2)
        return self._l.phieng_world_fluid_set_velocity(self.ptr, _fptr(v), len(v))

    def step(self, dt):
        self._l.phieng_world_step(self.ptr, float(dt))

    def sync(self):
        if not self._l.phieng_world_sync(self.ptr):
            raise RuntimeError(last_error())

    def set_record(self, flags):
        if not self._l.phieng_world_fluid_set_record(self.ptr, int(flags)):
            raise RuntimeError(last_error())

    def get_record(self):
        n = self.count()
        cells = np.zeros((n, 2), dtype=np.int32)
        order = np.zeros(n, dtype=np.uint32)
        rho = np.zeros(n, dtype=np.float32)
        p = np.zeros(n, dtype=np.float32)
        acc = np.zeros((n, 2), dtype=np.float32)
        cnt = np.zeros(n, dtype=np.uint32)
        got = self._l.phieng_world_fluid_get_record(
            self.ptr, cells.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            order.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _fptr(rho), _fptr(p), _fptr(acc),
            cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n)
        if got != n:
            raise RuntimeError(f"get_record returned {got}: {last_error()}")
        return dict(cells=cells, order=order, density=rho, pressure=p, accel=acc, count=cnt)

    def boundary_info(self):
        c, r = ctypes.c_uint32(), ctypes.c_uint64()
        self._l.phieng_world_boundary_info(self.ptr, ctypes.byref(c), ctypes.byref(r))
        return c.value, r.value

    def set_neighbour_list(self, verlet=True, skin=0.1):
        v = 2 if verlet == "auto" else (1 if verlet else 0)
        if not self._l.phieng_world_set_neighbour_list(self.ptr, v, float(skin)):
            raise ValueError("neighbour list settings rejected")

    def neighbour_stats(self):
        a, b, c, d = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        if not self._l.phieng_world_neighbour_stats(self.ptr, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d)):
            raise RuntimeError(last_error())
        return {"substeps": a.value, "rebuilds": b.value, "repaired": c.value, "list_substeps": d.value}

    def boundary(self):
        """Ghosts in generation order: dict(pos (G, 2), density, pressure)."""
        g, _ = self.boundary_info()
        xy = np.zeros((g, 2), np.float32)
        rho = np.zeros(g, np.float32)
        p = np.zeros(g, np.float32)
        got = self._l.phieng_world_boundary_get(self.ptr, _fptr(xy), _fptr(rho), _fptr(p), g)
        if got != g:
            raise RuntimeError(f"boundary_get returned {got} of {g}: {last_error()}")
        return dict(pos=xy, density=rho, pressure=p)


BATCH_KERNELS, BATCH_PERSISTENT, BATCH_PERSISTENT_GRAPH, BATCH_KERNELS_GRAPH = 0, 1, 2, 3


class Batch3:
    """Many worlds advanced together (``phieng_batch3_*``)."""

    def __init__(self, worlds):
        self._l = lib()
        self.worlds = list(worlds)
        arr = (ctypes.c_void_p * len(self.worlds))(*[w.ptr for w in self.worlds])
        self.ptr = self._l.phieng_batch3_new(arr, len(self.worlds))
        if not self.ptr:
            raise RuntimeError(self._l.phieng_batch3_last_error().decode())

    def close(self):
        if self.ptr:
            self._l.phieng_batch3_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, ok):
        if not ok:
            raise RuntimeError(self._l.phieng_batch3_last_error().decode())

    def set_mode(self, mode, blocks_per_cu=0):
        self._ck(self._l.phieng_batch3_set_mode(self.ptr, int(mode), int(blocks_per_cu)))

    def set_spin_limit(self, spins):
        self._ck(self._l.phieng_batch3_set_spin_limit(self.ptr, int(spins)))

    def step(self, dt):
        self._ck(self._l.phieng_batch3_step(self.ptr, float(dt)))

    def sync(self):
        self._ck(self._l.phieng_batch3_sync(self.ptr))

    def sync_worlds(self):
        self._ck(self._l.phieng_batch3_sync_worlds(self.ptr))

    def count(self):
        return self._l.phieng_batch3_count(self.ptr)

    def autotune(self, dt, trial_steps=10, cache_path=None):
        bpc = ctypes.c_uint32(0)
        ms = ctypes.c_double(0)
        m = self._l.phieng_batch3_autotune(self.ptr, float(dt), int(trial_steps),
                                           cache_path.encode() if cache_path else None,
                                           ctypes.byref(bpc), ctypes.byref(ms))
        if m < 0:
            raise RuntimeError(self._l.phieng_batch3_last_error().decode())
        return m, bpc.value, ms.value

    def set_timing(self, on):
        self._l.phieng_batch3_set_timing(self.ptr, 1 if on else 0)

    def get_timing(self):
        k = 32
        ms = (ctypes.c_double * k)()
        cnt = (ctypes.c_uint64 * k)()
        m = self._l.phieng_batch3_get_timing(self.ptr, ms, cnt, k)
        return {self._l.phieng_kernel_name(i).decode(): (ms[i], cnt[i]) for i in range(min(m, k)) if cnt[i]}


class Surrogate3:
    """The learned pressure term's network and training state (``phieng_surrogate3_*``)."""

    SAMPLE = 32

    def __init__(self, seed=0):
        self._l = lib()
        self.ptr = self._l.phieng_surrogate3_new(int(seed))
        if not self.ptr:
            raise RuntimeError(self._l.phieng_surrogate3_last_error().decode())
        self.nparam = self._l.phieng_surrogate3_num_params()

    def close(self):
        if self.ptr:
            self._l.phieng_surrogate3_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, ok):
        if not ok:
            raise RuntimeError(self._l.phieng_surrogate3_last_error().decode())

    def params(self):
        out = np.zeros(self.nparam, dtype=np.float32)
        self._ck(self._l.phieng_surrogate3_get_params(self.ptr, _fptr(out), self.nparam))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._ck(self._l.phieng_surrogate3_set_params(self.ptr, _fptr(p), len(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 16)
        y = np.zeros((len(x), 16), dtype=np.float32)
        self._ck(self._l.phieng_surrogate3_mlp_forward(self.ptr, _fptr(x), len(x), _fptr(y)))
        return y

    def backward(self, x, dy):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 16)
        dy = np.ascontiguousarray(dy, dtype=np.float32).reshape(-1, 16)
        g = np.zeros(self.nparam, dtype=np.float32)
        self._ck(self._l.phieng_surrogate3_mlp_backward(self.ptr, _fptr(x), _fptr(dy), len(x), _fptr(g)))
        return g

    def add_samples(self, s):
        s = np.ascontiguousarray(s, dtype=np.float32).reshape(-1, self.SAMPLE)
        self._ck(self._l.phieng_surrogate3_add_samples(self.ptr, _fptr(s), len(s)))

    def num_samples(self):
        return self._l.phieng_surrogate3_num_samples(self.ptr)

    def clear_samples(self):
        self._l.phieng_surrogate3_clear_samples(self.ptr)

    def samples(self, first=0, count=None):
        count = self.num_samples() - first if count is None else count
        out = np.zeros((count, self.SAMPLE), dtype=np.float32)
        self._ck(self._l.phieng_surrogate3_get_samples(self.ptr, _fptr(out), first, count))
        return out

    def train(self, epochs, batch=4096, lr=1e-3, seed=0):
        a, b, c = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        self._ck(self._l.phieng_surrogate3_train(self.ptr, epochs, batch, lr, seed, ctypes.byref(a),
                                                 ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def evaluate(self):
        a, b = ctypes.c_float(), ctypes.c_float()
        self._ck(self._l.phieng_surrogate3_eval(self.ptr, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

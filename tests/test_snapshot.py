"""Snapshots (phieng_world3_save / phieng_world3_load; ffi.rs:4036-4117,
crates/dynamics/src/snapshot.rs) without a GPU: worlds with bodies and a
fluid config but no particles.  The files are read back by independent
decoders (Python's json, struct and zlib.crc32) written from the
reference's encoder (encode_world3, :7920-8126; encode_envelope,
:5570-5579), and the reference's hostile-input test is restated."""
import json
import os
import struct
import zlib

import numpy as np

import pytest

from phi_eng_amd import ffi
from phi_eng_amd.ffi import FluidConfig3Ffi
from phi_eng_amd.world import World3


@pytest.fixture
def tmp(tmp_path):
    return lambda name: str(tmp_path / name)


def _world(cfg=None):
    w = World3()
    if cfg is not None:
        assert w.set_fluid_config(cfg)
    w.set_gravity((0.5, -9.81, 0.25))
    assert w.add_body((0.0, -1.0, 0.0), 1, 10.0, 0.5, 10.0, density=2.0) == 0
    assert w.add_body((1.0, 2.0, 3.0), 0, 0.5) == 1
    assert w.add_body((-1.0, 0.5, 0.0), 2, 0.25, 0.75) == 2
    return w


class Payload:
    """Little-endian reader of a payload, the order of encode_world3."""

    def __init__(self, b):
        self.b, self.p = b, 0

    def take(self, fmt):
        v = struct.unpack_from("<" + fmt, self.b, self.p)
        self.p += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]


def decode(path):
    b = open(path, "rb").read()
    assert b[:4] == b"PHE1"
    version, kind, n = struct.unpack_from("<IBI", b, 4)
    assert kind == 3 and len(b) == 13 + n + 4
    assert zlib.crc32(b[13:13 + n]) == struct.unpack_from("<I", b, 13 + n)[0]
    r = Payload(b[13:13 + n])
    out = dict(schema=version, gravity=r.take("3f"), dt=r.take("f"), sleep=r.take("2f"),
               solver=r.take("I3f"))
    if version >= 7:
        out["solver_flags"] = r.take("2B")
    out["validation"] = r.take("B5f2B")
    if version >= 4:
        out["gravity_config"] = r.take("II2f3f2f")
    if version >= 5:
        out["ambient"] = r.take("B3f3fB4f3fB3fIBB")
        out["time"] = r.take("f")
    if version >= 6:
        out["ccd"] = r.take("BBfI")
    out["next_body"], out["next_joint"], nb = r.take("3I")
    bodies = []
    for _ in range(nb):
        body = dict(id=r.take("I"), flags=r.take("B"), damping=r.take("2f"), material=r.take("3f"))
        tag = r.take("B")
        body["shape"] = (tag, r.take({0: "f", 1: "3f", 2: "2f"}[tag]))
        body["pos"], body["rot"], body["vel"], body["angvel"] = r.take("3f"), r.take("4f"), r.take("3f"), r.take("3f")
        if version >= 4:
            body["gravity_scale"] = r.take("f")
        if version >= 5:
            body["ambient_scale"] = r.take("f")
        bodies.append(body)
    out["bodies"] = bodies
    out["joints"], out["contacts"] = r.take("I"), r.take("I")
    cfg = dict(head=r.take("4f"))
    if version >= 5:
        cfg["opts"] = r.take("4fIBffBf")
        cfg["phases"] = (r.take("B"), r.take("4f"), r.take("4f"))
    cfg["tail"] = r.take("4fB")
    out["fluid_config"] = cfg
    out["particle_count"], out["particles"] = fluid_particles(r, version, 3)
    assert r.p == len(r.b), "trailing bytes"
    return out


MAX_PARTICLES_3D = 4096  # crates/dynamics/src/fluid3.rs:28; decode_fluid3 (snapshot.rs:6679-6684) refuses more


def fluid_particles(r, version, dim):
    """The particle block of encode_fluid3 / encode_fluid2 (snapshot.rs:
    6662-6677): a u32 count, then per particle dim position and dim velocity
    f32 values and, from schema 5 (SNAPSHOT_SCHEMA_AMBIENT), its mass (f32)
    and phase (u8), packed.  Returns (count, dict of arrays)."""
    n = r.take("I")
    fields = [("pos", "<f4", (dim,)), ("vel", "<f4", (dim,))]
    if version >= 5:
        fields += [("mass", "<f4"), ("phase", "u1")]
    dt = np.dtype(fields)  # packed: no padding between fields
    a = np.frombuffer(r.b, dtype=dt, count=n, offset=r.p)
    r.p += n * dt.itemsize
    out = dict(pos=np.array(a["pos"]).reshape(-1, dim), vel=np.array(a["vel"]).reshape(-1, dim))
    if version >= 5:
        out["mass"], out["phase"] = np.array(a["mass"]), np.array(a["phase"])
    return n, out


def json_particles(d, dim):
    """The same fields from the JSON form (FluidParticle3Dto / 2Dto,
    snapshot.rs:582-592: position and velocity objects, mass, phase)."""
    ps = d["fluid"]["particles"]
    ax = "xyz"[:dim]
    pos = np.array([[p["position"][c] for c in ax] for p in ps], np.float64).astype(np.float32)
    vel = np.array([[p["velocity"][c] for c in ax] for p in ps], np.float64).astype(np.float32)
    mass = np.array([p.get("mass", 0.0) for p in ps], np.float64).astype(np.float32)
    phase = np.array([p.get("phase", 0) for p in ps], np.uint8)
    return dict(pos=pos.reshape(-1, dim), vel=vel.reshape(-1, dim), mass=mass, phase=phase)


def test_default_world_json_layout(tmp):
    """A default-config world is written at schema 3 with the reference's
    field names and defaults (World3::new, RigidBody3::new, the sanitised
    material of the C ABI's add_body)."""
    w = _world()
    assert w.save(tmp("w.json")) == 1
    d = json.load(open(tmp("w.json")))
    assert d["schema"] == 3
    assert d["gravity"] == {"x": 0.5, "y": -9.81, "z": 0.25}
    assert abs(d["dt"] - 1 / 60) < 1e-7 and d["linear_sleep_threshold"] == 0.05
    assert d["solver"] == {"iterations": 8, "slop": 0.005, "baumgarte": 0.2, "early_exit_tolerance": 1e-06}
    assert d["validation"]["max_coordinate"] == 10000 and d["validation"]["auto_correct"] is True
    assert d["next_body_id"] == 3 and d["joints"] == [] and d["contacts"] == []
    b0, b1, b2 = d["bodies"]
    assert b0["shape"] == {"type": "box", "half_extents": {"x": 10, "y": 0.5, "z": 10}}
    assert b1["shape"] == {"type": "sphere", "radius": 0.5}
    assert b2["shape"] == {"type": "capsule", "radius": 0.25, "half_length": 0.75}
    assert b0["material"]["density"] == 2 and b0["rotation"] == [0, 0, 0, 1] and b0["is_static"]
    assert b0["linear_damping"] == 0.01 and "gravity_scale" not in b0
    c = d["fluid"]["config"]
    assert set(c) == {"particle_mass", "rest_density", "smoothing_length", "stiffness", "viscosity", "drag",
                      "max_velocity", "particle_radius", "enabled"}
    assert d["fluid"]["particles"] == []
    w.close()


def test_binary_layout_and_schema_5(tmp):
    """The binary envelope and payload follow encode_world3; any fluid
    option lifts the schema to 5 (gravity and ambient blocks, per-body
    scales, the long fluid config with the phase arrays)."""
    w = _world()
    assert w.save(tmp("a.bin"), binary=True) == 1
    a = decode(tmp("a.bin"))
    assert a["schema"] == 3 and a["next_body"] == 3 and a["joints"] == 0 and a["contacts"] == 0
    assert [b["shape"][0] for b in a["bodies"]] == [1, 0, 2]
    assert a["bodies"][0]["flags"] == 1 and a["bodies"][0]["rot"] == (0.0, 0.0, 0.0, 1.0)
    assert a["particle_count"] == 0
    cfg = FluidConfig3Ffi.default()
    cfg.gamma = 7.0
    cfg.phase_count = 2
    cfg.phase_rest_density[1] = 2.0
    cfg.phase_mass[1] = 0.125
    w2 = _world(cfg)
    assert w2.save(tmp("b.bin"), binary=True) == 1
    b = decode(tmp("b.bin"))
    assert b["schema"] == 5
    assert b["gravity_config"] == (0, 0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.5)
    assert b["ambient"][0] == 0 and b["time"] == 0.0
    assert all(x["gravity_scale"] == 1.0 and x["ambient_scale"] == 1.0 for x in b["bodies"])
    assert b["fluid_config"]["opts"][0] == 7.0
    assert b["fluid_config"]["phases"] == (2, (1.0, 2.0, 1.0, 1.0), (0.0625, 0.125, 0.0625, 0.0625))
    assert w2.save(tmp("b.json")) == 1
    d = json.load(open(tmp("b.json")))
    assert d["schema"] == 5 and d["fluid"]["config"]["gamma"] == 7
    assert d["fluid"]["config"]["phases"] == {"count": 2, "rest_density": [1, 2, 1, 1], "mass": [0.0625, 0.125,
                                                                                               0.0625, 0.0625]}
    w.close()
    w2.close()


@pytest.mark.parametrize("binary", [False, True])
def test_round_trip_is_a_fixed_point(tmp, binary):
    """save(load(save(w))) == save(w), byte for byte; the loaded world keeps
    gravity, bodies and config."""
    cfg = FluidConfig3Ffi.default()
    cfg.cfl = 0.3
    cfg.max_substeps = 4
    cfg.boundary_particles = 1
    w = _world(cfg)
    ext = "bin" if binary else "json"
    assert w.save(tmp("1." + ext), binary=binary) == 1
    v = World3.load(tmp("1." + ext))
    assert v is not None, ffi.last_error()
    assert bytes(v.get_fluid_config()) == bytes(w.get_fluid_config())
    assert v.save(tmp("2." + ext), binary=binary) == 1
    assert open(tmp("1." + ext), "rb").read() == open(tmp("2." + ext), "rb").read()
    w.close()
    v.close()


def test_load_applies_reference_rules(tmp):
    """apply_fluid3 (snapshot.rs:4562-4584): a single-phase file loads with
    phase rest densities 1 and phase masses = particle_mass; a missing fluid
    block is FluidConfig3::default; the solver switches (schema 7) and the
    CCD block (6) survive a round trip."""
    w = _world()
    assert w.save(tmp("w.json")) == 1
    d = json.load(open(tmp("w.json")))
    d["fluid"]["config"]["rest_density"] = 1000.0
    d["fluid"]["config"]["particle_mass"] = 0.5
    d["solver"]["friction"] = False
    json.dump(d, open(tmp("m.json"), "w"))
    v = World3.load(tmp("m.json"))
    assert v is not None, ffi.last_error()
    c = v.get_fluid_config()
    assert c.rest_density == 1000.0 and list(c.phase_rest_density) == [1.0] * 4
    assert list(c.phase_mass) == [0.5] * 4
    assert v.save(tmp("m.bin"), binary=True) == 1 and decode(tmp("m.bin"))["schema"] == 7
    assert decode(tmp("m.bin"))["solver_flags"] == (0, 1)
    del d["fluid"]
    d["ccd"] = {"enabled": True, "tier": 1, "max_movement": 0.2, "max_substeps": 4}
    del d["solver"]["friction"]
    json.dump(d, open(tmp("n.json"), "w"))
    v2 = World3.load(tmp("n.json"))
    assert v2 is not None, ffi.last_error()
    assert bytes(v2.get_fluid_config()) == bytes(FluidConfig3Ffi.default())
    assert v2.save(tmp("n.bin"), binary=True) == 1
    assert decode(tmp("n.bin"))["ccd"] == (1, 1, pytest.approx(0.2), 4)
    w.close()
    v.close()
    v2.close()


def test_hostile_input_rejected(tmp):
    """The reference's e05 (crates/engine/tests/snapshot_tests.rs:115-149)
    on the 3D loader, plus what the GPU path cannot hold."""
    w = _world()
    assert w.save(tmp("ok.bin"), binary=True) == 1
    assert w.save(tmp("ok.json")) == 1
    good = open(tmp("ok.bin"), "rb").read()

    def rejected(data, name="x"):
        open(tmp(name), "wb").write(data if isinstance(data, bytes) else data.encode())
        return World3.load(tmp(name)) is None

    assert rejected("{nope")
    assert rejected(b"junk")
    v = bytearray(good)
    v[4] = 99
    n = struct.unpack_from("<I", v, 9)[0]
    struct.pack_into("<I", v, 13 + n, zlib.crc32(bytes(v[13:13 + n])))
    assert rejected(bytes(v)) and "version" in ffi.last_error()
    c = bytearray(good)
    c[20] ^= 0xFF
    assert rejected(bytes(c)) and "CRC" in ffi.last_error()
    assert rejected(good[:len(good) // 2])
    v = bytearray(good)
    v[4] = 9
    struct.pack_into("<I", v, 13 + n, zlib.crc32(bytes(v[13:13 + n])))
    assert rejected(bytes(v)) and "schema 9" in ffi.last_error()
    d = json.load(open(tmp("ok.json")))
    for mutate, why in [
        (lambda d: d["bodies"][1].update(velocity={"x": 1.0, "y": 0.0, "z": 0.0}), "moving static"),
        (lambda d: d["bodies"][1].update(rotation=[0, 0.6, 0, 0.8]), "rotated"),
        (lambda d: d["bodies"][1].update(id=5), "out of order"),
        (lambda d: d.update(next_body_id=7), "next_body_id"),
        (lambda d: d.update(joints=[{"type": "distance3"}]), "joints"),
        (lambda d: d.pop("contacts"), "contacts"),
        (lambda d: d["fluid"]["config"].update(gamma=0.5), "validation"),
        (lambda d: d.update(schema=2), "version"),
    ]:
        e = json.loads(json.dumps(d))
        mutate(e)
        assert rejected(json.dumps(e), "h.json"), why
        assert why.split()[0] in ffi.last_error() or why == "out of order" and "order" in ffi.last_error(), \
            (why, ffi.last_error())
    w.close()


def test_argument_checks(tmp):
    """ffi.rs:4050-4070: null world or path, bad length, mode or format
    return 0 / null; an unwritable path returns -1."""
    lib = ffi.lib()
    w = _world()
    p = tmp("a.json").encode()
    assert lib.phieng_world3_save(None, p, len(p), 0, 0) == 0
    assert lib.phieng_world3_save(w.ptr, None, 5, 0, 0) == 0
    assert lib.phieng_world3_save(w.ptr, p, 0, 0, 0) == 0
    assert lib.phieng_world3_save(w.ptr, p, 5000, 0, 0) == 0
    assert lib.phieng_world3_save(w.ptr, p, len(p), 2, 0) == 0
    assert lib.phieng_world3_save(w.ptr, p, len(p), 0, 2) == 0
    assert lib.phieng_world3_save(w.ptr, b"\xff\xfe", 2, 0, 0) == 0
    d = os.path.dirname(tmp("a.json")).encode()
    assert lib.phieng_world3_save(w.ptr, d, len(d), 0, 0) == -1
    assert lib.phieng_world3_save(w.ptr, p, len(p), 0, 1) == 1
    assert lib.phieng_world3_load(None, 3, 0) is None
    assert lib.phieng_world3_load(p, len(p), 7) is None
    q = tmp("missing.json").encode()
    assert lib.phieng_world3_load(q, len(q), 0) is None
    w.close()


# ---------------------------------------------------------------- 2D (World2Dto)

def _world2(cfg=None):
    from phi_eng_amd.world import SHAPE2_BOX, SHAPE2_CIRCLE, World2
    w = World2()
    if cfg is not None:
        assert w.set_fluid_config(cfg)
    w.set_gravity((0.5, -9.81))
    assert w.add_body((0.0, -1.0), SHAPE2_BOX, 10.0, 0.5, density=2.0) == 0
    assert w.add_body((1.0, 2.0), SHAPE2_CIRCLE, 0.5) == 1
    return w


def decode2(path):
    """encode_world2 (snapshot.rs:6033-6238) read back with struct."""
    b = open(path, "rb").read()
    assert b[:4] == b"PHE1"
    version, kind, n = struct.unpack_from("<IBI", b, 4)
    assert kind == 2 and len(b) == 13 + n + 4
    assert zlib.crc32(b[13:13 + n]) == struct.unpack_from("<I", b, 13 + n)[0]
    r = Payload(b[13:13 + n])
    out = dict(schema=version, gravity=r.take("2f"), dt=r.take("f"), sleep=r.take("2f"), solver=r.take("I3f"))
    out["validation"] = r.take("B5f2B")
    if version >= 4:
        out["gravity_config"] = r.take("II2f2f2f")
    if version >= 5:
        out["ambient"] = r.take("B2f3fB4f2fB3fIBB")
        out["time"] = r.take("f")
    out["next_body"], out["next_joint"], out["cell_size"], nb = r.take("IIfI")
    bodies = []
    for _ in range(nb):
        body = dict(id=r.take("I"), flags=r.take("B"), damping=r.take("2f"), material=r.take("3f"))
        tag = r.take("B")
        body["shape"] = (tag, r.take({0: "f", 1: "2f"}[tag]))
        body["pos"], body["rot"], body["vel"], body["angvel"] = r.take("2f"), r.take("4f"), r.take("2f"), r.take("f")
        if version >= 4:
            body["gravity_scale"] = r.take("f")
        if version >= 5:
            body["ambient_scale"] = r.take("f")
        bodies.append(body)
    out["bodies"] = bodies
    out["joints"], out["contacts"] = r.take("I"), r.take("I")
    cfg = dict(head=r.take("4f"))
    if version >= 5:
        cfg["opts"] = r.take("4fIBffBf")
        cfg["phases"] = (r.take("B"), r.take("4f"), r.take("4f"))
    cfg["tail"] = r.take("4fB")
    out["fluid_config"] = cfg
    out["particle_count"], out["particles"] = fluid_particles(r, version, 2)
    assert r.p == len(r.b), "trailing bytes"
    return out


def test_2d_layouts_and_round_trip(tmp):
    """World2Dto in both formats: 2-vectors, cell_size, circle and box
    shapes, the identity rotation matrix, a scalar angular velocity; the
    round trip is a fixed point; a 3D file is refused by the 2D loader."""
    from phi_eng_amd.world import World2
    cfg = FluidConfig3Ffi.default()
    cfg.wall_friction = 0.25
    w = _world2(cfg)
    assert w.save(tmp("w2.json")) == 1, ffi.last_error()
    d = json.load(open(tmp("w2.json")))
    assert d["schema"] == 5 and d["gravity"] == {"x": 0.5, "y": -9.81} and d["cell_size"] == 1
    b0, b1 = d["bodies"]
    assert b0["shape"] == {"type": "box", "half_extents": {"x": 10, "y": 0.5}}
    assert b1["shape"] == {"type": "circle", "radius": 0.5}
    assert b0["rotation"] == [1, 0, 0, 1] and b0["angular_velocity"] == 0 and b0["position"] == {"x": 0, "y": -1}
    assert d["fluid"]["config"]["wall_friction"] == 0.25
    assert w.save(tmp("w2.bin"), binary=True) == 1
    e = decode2(tmp("w2.bin"))
    assert e["schema"] == 5 and e["cell_size"] == 1.0 and e["next_body"] == 2
    assert [x["shape"] for x in e["bodies"]] == [(1, (10.0, 0.5)), (0, 0.5)]
    assert e["bodies"][0]["rot"] == (1.0, 0.0, 0.0, 1.0)
    for ext, binary in (("json", False), ("bin", True)):
        v = World2.load(tmp("w2." + ext))
        assert v is not None, ffi.last_error()
        assert bytes(v.get_fluid_config()) == bytes(w.get_fluid_config())
        assert v.save(tmp("v2." + ext), binary=binary) == 1
        assert open(tmp("w2." + ext), "rb").read() == open(tmp("v2." + ext), "rb").read()
        v.close()
    w3 = _world()
    assert w3.save(tmp("w3.bin"), binary=True) == 1
    assert World2.load(tmp("w3.bin")) is None and "kind" in ffi.last_error()
    assert World3.load(tmp("w2.bin")) is None and "kind" in ffi.last_error()
    d["bodies"][1]["shape"] = {"type": "polygon", "vertices": [{"x": 0, "y": 0}]}
    json.dump(d, open(tmp("p.json"), "w"))
    assert World2.load(tmp("p.json")) is None and "polygon" in ffi.last_error()
    w.close()
    w3.close()


@pytest.mark.parametrize("binary", [False, True])
def test_dynamic_body_state_round_trip(tmp, binary):
    """A dynamic body keeps its state through save / load (body3_dto,
    load_body3: is_static false, position, rotation, velocities); the
    loaded world steps on exactly as the saved one (host-only steps of an
    empty fluid), and save(load(save(w))) == save(w)."""
    w = _world()
    b = w.add_body((0.5, 3.0, -0.5), 1, 0.2, 0.3, 0.4, density=1.5, is_static=0)
    dt = np.float32(1.0 / 60.0)
    for _ in range(7):
        w.step(dt)
    ext = "bin" if binary else "json"
    assert w.save(tmp("d1." + ext), binary=binary) == 1, ffi.last_error()
    if not binary:
        d = json.load(open(tmp("d1." + ext)))
        jb = d["bodies"][b]
        assert jb["is_static"] is False
        assert abs(jb["velocity"]["y"] - float(w.body_velocity(b)[0][1])) < 1e-6
    v = World3.load(tmp("d1." + ext))
    assert v is not None, ffi.last_error()
    for x, y in ((w.body_transform(b), v.body_transform(b)), (w.body_velocity(b), v.body_velocity(b))):
        for a1, a2 in zip(x, y):
            assert np.array_equal(a1.view(np.uint32), a2.view(np.uint32))
    assert v.save(tmp("d2." + ext), binary=binary) == 1
    assert open(tmp("d1." + ext), "rb").read() == open(tmp("d2." + ext), "rb").read()
    for _ in range(5):
        w.step(dt)
        v.step(dt)
    assert np.array_equal(w.body_transform(b)[0].view(np.uint32), v.body_transform(b)[0].view(np.uint32))
    w.close()
    v.close()

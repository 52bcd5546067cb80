"""Dynamic bodies coupled both ways to the fluid (fluid3.rs:1141-1213,
body.rs:1334-1406) through the C ABI, against the oracle's world step
(oracle3_world_run_f32, Jacobi integrate): particles and bodies bit for
bit, step after step, and the reference's own box-floats test
(fluid3.rs:1349-1396)."""
import numpy as np
import pytest

from phi_eng_amd.ffi import FluidConfig3Ffi
from phi_eng_amd.world import World3

pytestmark = pytest.mark.gpu

TANK = [((2.0, -1.0, 2.0), (3.0, 0.5, 3.0)), ((-1.0, 2.0, 2.0), (0.5, 3.5, 3.5)),
        ((5.0, 2.0, 2.0), (0.5, 3.5, 3.5)), ((2.0, 2.0, -1.0), (3.5, 3.5, 0.5)),
        ((2.0, 2.0, 5.0), (3.5, 3.5, 0.5))]
POOL = [((0.375, 1.0, 1.5), (0.125, 2.0, 1.25)), ((2.625, 1.0, 1.5), (0.125, 2.0, 1.25)),
        ((1.5, 1.0, 0.375), (1.0, 2.0, 0.125)), ((1.5, 1.0, 2.625), (1.0, 2.0, 0.125))]


def _lattice(origin, nx, ny, nz, dx):
    """fluid3.rs:1251-1266 (f32 arithmetic)."""
    f = np.float32
    pts = [(f(origin[0]) + f(ix) * f(dx), f(origin[1]) + f(iy) * f(dx), f(origin[2]) + f(iz) * f(dx))
           for iz in range(nz) for iy in range(ny) for ix in range(nx)]
    return np.array(pts, np.float32)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


class Pair:
    """The GPU world and the oracle's, advanced together."""

    def __init__(self, O, cfg, statics, pos):
        self.O = O
        self.cfg = cfg
        self.w = World3()
        assert self.w.set_fluid_config(cfg)
        self.bodies = []
        for c, he in statics:
            assert self.w.add_static_box(c, he) >= 0
            self.bodies.append(O.rigid(1, c, *he))
        assert self.w.add_particles(pos) == len(pos)
        self.pos = pos.copy()
        self.vel = np.zeros_like(pos)

    def add_dynamic(self, shape, center, a, b=0.0, c=0.0, density=1.0):
        bid = self.w.add_body(center, shape, a, b, c, density=density, is_static=0)
        assert bid == len(self.bodies), bid
        self.bodies.append(self.O.rigid(shape, center, a, b, c, density=density, is_static=False))
        return bid

    def step(self, dt, k):
        self.w.step(dt)
        self.pos, self.vel, _ = self.O.world_run_f32(self.pos, self.vel, self.cfg, (0.0, -9.81, 0.0),
                                                     self.bodies, dt, 1, jacobi=True)
        p, v = self.w.get_state()
        assert np.isfinite(self.pos).all() and np.isfinite(self.vel).all(), f"the scene blew up at step {k}"
        assert np.array_equal(_bits(p), _bits(self.pos)), f"positions differ at step {k}"
        assert np.array_equal(_bits(v), _bits(self.vel)), f"velocities differ at step {k}"
        for bid, b in enumerate(self.bodies):
            if b.is_static:
                continue
            tp, tq = self.w.body_transform(bid)
            bv, bw = self.w.body_velocity(bid)
            for got, ref, what in ((tp, b.pos, "position"), (tq, b.rot, "rotation"), (bv, b.vel, "velocity"),
                                   (bw, b.ang, "angular velocity")):
                assert np.array_equal(_bits(got), _bits(np.array(ref[:], np.float32))), \
                    f"body {bid} {what} at step {k}: {got} vs {list(ref)}"


def test_box_floats(gpu, oracle_lib):
    """fluid3.rs:1349-1396: 200 settling steps of a 6 x 8 x 6 pool, then a
    half-density box placed above the surface floats (0 < y < 2) for 400
    steps; every step equals the oracle bit for bit."""
    cfg = FluidConfig3Ffi.default()
    pr = Pair(oracle_lib, cfg, TANK + POOL, _lattice((0.75, -0.5, 0.75), 6, 8, 6, 0.25))
    dt = np.float32(1.0 / 60.0)
    for k in range(200):
        pr.step(dt, k)
    b = pr.add_dynamic(1, (1.5, 1.4, 1.5), 0.3, 0.3, 0.3, density=0.5)
    for k in range(400):
        pr.step(dt, 200 + k)
        if k % 100 == 0:
            print(f"[box floats] step {k} box y {pr.w.body_transform(b)[0][1]:.4f}", flush=True)
    y = pr.w.body_transform(b)[0][1]
    p, v = pr.w.get_state()
    assert np.isfinite(p).all() and np.isfinite(v).all()
    assert 0.0 < y < 2.0, y
    q = pr.w.body_transform(b)[1]
    assert not np.array_equal(q, [0, 0, 0, 1]), "the box never turned: the rotated paths were not exercised"
    pr.w.close()


def test_sphere_and_capsule_with_wall_friction(gpu, oracle_lib):
    """A sphere and a capsule dropped into a sloshing pool with wall
    friction (the friction reads each body's velocity as the earlier
    particles' impulses left it): bit for bit over 150 steps."""
    cfg = FluidConfig3Ffi.default()
    cfg.wall_friction = 0.3
    pos = _lattice((0.75, -0.5, 0.75), 7, 6, 7, 0.22)
    pr = Pair(oracle_lib, cfg, TANK + POOL, pos)
    pr.vel[:, 0] = 0.5
    pr.w.set_velocity(pr.vel)
    pr.add_dynamic(0, (1.2, 1.3, 1.3), 0.3, density=3.0)
    pr.add_dynamic(2, (1.8, 1.6, 1.7), 0.2, 0.3, density=3.0)
    dt = np.float32(1.0 / 60.0)
    for k in range(150):
        pr.step(dt, k)
    pr.w.close()


def test_free_fall_without_fluid(gpu, oracle_lib):
    """No particles: the body still falls (gravity kick and integrate)."""
    cfg = FluidConfig3Ffi.default()
    w = World3()
    assert w.set_fluid_config(cfg)
    bid = w.add_body((0.0, 5.0, 0.0), 0, 0.5, density=2.0, is_static=0)
    ref = [oracle_lib.rigid(0, (0.0, 5.0, 0.0), 0.5, density=2.0, is_static=False)]
    dt = np.float32(1.0 / 60.0)
    for k in range(30):
        w.step(dt)
        oracle_lib.world_run_f32(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), cfg,
                                 (0.0, -9.81, 0.0), ref, dt, 1)
    p, _ = w.body_transform(bid)
    v, _ = w.body_velocity(bid)
    assert np.array_equal(_bits(p), _bits(np.array(ref[0].pos[:], np.float32)))
    assert np.array_equal(_bits(v), _bits(np.array(ref[0].vel[:], np.float32)))
    assert v[1] < -4.0
    w.close()

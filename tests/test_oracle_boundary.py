"""Pins of the oracle's boundary ghosts (config boundary_particles):
oracle/sph_oracle.cpp and oracle/sph2d_oracle.cpp restating
crates/dynamics/src/fluid3.rs:660-920, 1018-1027, 1074-1084 and
crates/dynamics/src/fluid.rs:617-818, 929-995.

Ghost positions are checked against values worked out by hand from the
generation rules (so a changed face layout, count or offset fails), the
ghost density against brute force, the ghost force against the two-body
closed form (one ghost, one particle: a wrong sign or a dropped pressure
term fails), and the reference's own tests are restated.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from phi_eng_amd.ffi import FluidConfig3Ffi


def _cfg3(h=0.5, **kw):
    c = FluidConfig3Ffi.default()
    c.smoothing_length = h
    c.boundary_particles = 1
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _ghosts3(bodies, cfg=None, pos=None, mass=None):
    cfg = cfg or _cfg3()
    pos = np.zeros((1, 3), np.float32) + 1e6 if pos is None else pos
    mass = np.full(len(pos), 0.0625, np.float32) if mass is None else mass
    return O.boundary_f32(pos, mass, cfg, bodies)


def test_box_ghosts_known_answer(oracle_lib):
    """A cube of half extent 0.25 at (1, 2, 3), h = 0.5 (spacing 0.25):
    area 0.25 per face -> 4 ghosts per sweep, t = 1/8, 3/8, 5/8, 7/8,
    (2t - 1) he = -3/16, -1/16, 1/16, 3/16 and (2 ((7t) mod 1) - 1) he =
    3/16, 1/16, -1/16, -3/16.  The reference leaves the first two sweep
    pairs on the mid-planes x = 0 and y = 0 and puts the third pair at
    z = -he (fluid3.rs:751-783)."""
    g = _ghosts3(O.bodies_from_boxes([((1.0, 2.0, 3.0), (0.25, 0.25, 0.25))]))["pos"]
    assert g.shape == (24, 3)
    u = np.array([-3, -1, 1, 3], np.float64) / 16
    w = -u
    want = []
    for axis in (0, 0, 1, 1, 2, 2):
        for k in range(4):
            if axis == 0:
                want.append((0.0, u[k], w[k]))
            elif axis == 1:
                want.append((u[k], 0.0, w[k]))
            else:
                want.append((u[k], w[k], -0.25))
    want = np.array(want) + [1.0, 2.0, 3.0]
    assert np.abs(g - want).max() <= 1e-6


def test_sphere_ghosts_known_answer(oracle_lib):
    """Radius 0.25, spacing 0.25: ceil(4 pi) = 13 Fibonacci points on the
    sphere; ghost 0 = R (5/13, 12/13, 0); the y levels step by 2R/13."""
    c = np.array([0.5, -1.0, 2.0])
    g = _ghosts3(O.bodies_from_boxes([], spheres=[(tuple(c), 0.25)]))["pos"].astype(np.float64)
    assert g.shape == (13, 3)
    R = 0.25
    assert np.abs(g[0] - (c + R * np.array([5 / 13, 12 / 13, 0.0]))).max() <= 1e-6
    assert np.abs(np.linalg.norm(g - c, axis=1) - R).max() <= 1e-6
    assert np.abs(np.diff(g[:, 1]) + 2 * R / 13).max() <= 1e-6
    golden = math.pi * (3 - math.sqrt(5))
    ang = np.arctan2(g[:, 2] - c[2], g[:, 0] - c[0])
    d = np.angle(np.exp(1j * (ang - golden * np.arange(13))))
    assert np.abs(d).max() <= 1e-5


def test_capsule_ghosts_known_answer(oracle_lib):
    """Radius 0.1, half length 0.2, spacing 0.25: ceil(4 pi 0.02 / 0.0625) = 5
    points per ring, rings at y = -0.1 and 0.1, then the ends at -/+ 0.3."""
    g = _ghosts3(O.bodies_from_boxes([], capsules=[((0.0, 0.0, 0.0), 0.1, 0.2)]))["pos"].astype(np.float64)
    assert g.shape == (12, 3)
    a = 2 * math.pi * np.arange(5) / 5
    for ring, y in ((0, -0.1), (1, 0.1)):
        want = np.stack([0.1 * np.cos(a), np.full(5, y), 0.1 * np.sin(a)], axis=1)
        assert np.abs(g[5 * ring:5 * ring + 5] - want).max() <= 1e-6
    assert np.abs(g[10:] - [[0, -0.3, 0], [0, 0.3, 0]]).max() <= 1e-6


def test_reference_tank_ghost_count(oracle_lib):
    """The reference's test tank (fluid3.rs:1229-1248), h = 0.5: face sweeps
    of ceil(area / 0.0625) clamped to 512 — floor 2 (96 + 512 + 96), each
    wall 2 (512 + 112 + 112)."""
    tank = [((2.0, -1.0, 2.0), (3.0, 0.5, 3.0)), ((-1.0, 2.0, 2.0), (0.5, 3.5, 3.5)),
            ((5.0, 2.0, 2.0), (0.5, 3.5, 3.5)), ((2.0, 2.0, -1.0), (3.5, 3.5, 0.5)),
            ((2.0, 2.0, 5.0), (3.5, 3.5, 0.5))]
    g = _ghosts3(O.bodies_from_boxes(tank))
    assert len(g["pos"]) == 2 * (96 + 512 + 96) + 4 * 2 * (512 + 112 + 112)


def test_ghost_density_brute_force(oracle_lib):
    """Ghost density = sum_j m_j W(|x_j - g|) over the fluid within h
    (clamped to 3 rho0), pressure = EOS of the unclamped sum with the base
    rest density (fluid3.rs:906-920)."""
    rng = np.random.default_rng(3)
    h = 0.1
    cfg = _cfg3(h=h, rest_density=1000.0, stiffness=50.0, gamma=7.0)
    pos = (rng.random((900, 3)) * [0.6, 0.3, 0.6] + [-0.3, 0.0, -0.3]).astype(np.float32)
    mass = np.full(900, 1.7, np.float32)
    bodies = O.bodies_from_boxes([((0.0, -0.05, 0.0), (0.25, 0.05, 0.25))])
    out = O.boundary_f32(pos, mass, cfg, bodies)
    g = out["pos"].astype(np.float64)
    assert len(g) > 50
    d2 = ((pos[None, :, :].astype(np.float64) - g[:, None, :]) ** 2).sum(axis=2)
    w = np.where(d2 < h * h, 315 / (64 * math.pi * h ** 9) * np.maximum(h * h - d2, 0) ** 3, 0.0)
    rho = (w * mass[None, :]).sum(axis=1)
    assert (rho > 0).any()
    assert np.abs(out["density"] - np.minimum(rho, 3000.0)).max() <= 2e-5 * rho.max()
    p = np.maximum(50.0 * 1000.0 * ((rho / 1000.0) ** 7 - 1.0), 0.0)
    assert np.abs(out["pressure"] - p).max() <= 1e-4 * max(p.max(), 1.0)


def test_single_ghost_force_closed_form(oracle_lib):
    """A sphere small enough for one ghost (at centre + (R, 0, 0)) and one
    particle at distance d: rho_b = m W(d), rho_i = m W(0),
    a = g - grad W(x_i - x_b) m_b (p_i / rho_i^2 + p_b / rho_b^2) / rho_i."""
    h, m, k, rest = 0.5, 0.0625, 100.0, 0.5  # no density clamp (3 rest > m W(0)), both pressures > 0
    cfg = _cfg3(h=h, stiffness=k, rest_density=rest, particle_mass=0.03, viscosity=0.0)
    for ph in range(4):
        cfg.phase_rest_density[ph] = rest
    R = 1e-3
    ghost = np.array([R, 0.0, 0.0])
    x = np.array([[R + 0.08, 0.06, 0.0]], np.float32)
    bodies = O.bodies_from_boxes([], spheres=[((0.0, 0.0, 0.0), R)])
    out = O.substep_f32(x, np.zeros_like(x), cfg, (0.0, 0.0, 0.0), bodies, 1e-4, jacobi=True)
    W = lambda r2: 315 / (64 * math.pi * h ** 9) * (h * h - r2) ** 3
    d = x[0].astype(np.float64) - ghost
    r = np.linalg.norm(d)
    rho_b, rho_i = m * W(r * r), m * W(0.0)
    assert rho_i < 3 * rest and rho_b > rest
    p_b, p_i = k * (rho_b - rest), k * (rho_i - rest)
    grad = d / r * (-45 / (math.pi * h ** 6) * (h - r) ** 2)
    a = -grad * 0.03 * (p_i / rho_i ** 2 + p_b / rho_b ** 2) / rho_i
    assert np.abs(out["accel"][0] - a).max() <= 1e-5 * np.linalg.norm(a)
    assert a[0] > 0  # pushed away from the ghost
    # the same particle without ghosts feels nothing
    cfg.boundary_particles = 0
    out0 = O.substep_f32(x, np.zeros_like(x), cfg, (0.0, 0.0, 0.0), bodies, 1e-4, jacobi=True)
    assert np.abs(out0["accel"]).max() == 0.0


def _tank3():
    return O.bodies_from_boxes([((2.0, -1.0, 2.0), (3.0, 0.5, 3.0)), ((-1.0, 2.0, 2.0), (0.5, 3.5, 3.5)),
                                ((5.0, 2.0, 2.0), (0.5, 3.5, 3.5)), ((2.0, 2.0, -1.0), (3.5, 3.5, 0.5)),
                                ((2.0, 2.0, 5.0), (3.5, 3.5, 0.5))])


def test_reference_boundary_contained_3d(oracle_lib):
    """fluid3.rs:1489-1518: a 4 x 6 x 4 lattice at 0.25 in the tank, default
    config with ghosts, 200 steps at 60 Hz: finite, contained, calm."""
    pos = np.array([(ix * 0.25, iy * 0.25, iz * 0.25) for iz in range(4) for iy in range(6) for ix in range(4)],
                   np.float32)
    cfg = _cfg3()
    p, v = pos, np.zeros_like(pos)
    peak = 0.0
    for _ in range(20):  # 20 x 10 steps, checking in between
        p, v, _ = O.run_f32(p, v, cfg, (0.0, -9.81, 0.0), _tank3(), 1.0 / 60.0, 10)
        assert np.isfinite(p).all() and np.isfinite(v).all()
        peak = max(peak, float(np.linalg.norm(v, axis=1).max()))
    assert (p[:, 0] > -1).all() and (p[:, 0] < 5).all() and (p[:, 2] > -1).all() and (p[:, 2] < 5).all()
    assert peak < 30.0
    assert len(_ghosts3(_tank3())["pos"]) > 0


# ---------------------------------------------------------------- 2D

def _cfg2(h=0.5):
    c = O.Oracle2Cfg.default()
    c.smoothing_length = h
    c.boundary_particles = 1
    return c


def test_2d_ghosts_known_answer(oracle_lib):
    """fluid.rs:673-714 at h = 0.5 (spacing 0.25): a circle of radius 0.1
    gives ceil(2 pi 0.1 / 0.25) = 3 ring points; a box of half extents
    (0.25, 0.125) gives edges of ceil(0.25 / 0.25) = 1 point along y (the
    first two, from (-a, -b) and (-a, b)) and ceil(0.5 / 0.25) = 2 points
    along x (from (-a, -b) and (a, -b)), at t = 1/2 resp. 1/4, 3/4."""
    b = O.bodies2(boxes=[((1.0, 1.0), (0.25, 0.125))], circles=[((0.0, 0.0), 0.1)])
    out = O.boundary2_f32(np.full((1, 2), 1e6, np.float32), np.full(1, 0.0625, np.float32), _cfg2(), b)
    g = out["pos"].astype(np.float64)
    # d = 2 extent t: 0.125 along y; 0.125 and 0.375 along x
    want_box = [(-0.25, 0.0), (-0.25, 0.25), (-0.125, -0.125), (0.125, -0.125), (0.375, -0.125), (0.625, -0.125)]
    a = 2 * math.pi * np.arange(3) / 3
    want = np.concatenate([np.array(want_box) + 1.0, np.stack([0.1 * np.cos(a), 0.1 * np.sin(a)], axis=1)])
    assert g.shape == want.shape
    assert np.abs(g - want).max() <= 1e-6


def test_2d_ghost_density_brute_force(oracle_lib):
    rng = np.random.default_rng(4)
    h = 0.1
    cfg = _cfg2(h)
    cfg.rest_density = 1000.0
    cfg.stiffness = 30.0
    cfg.gamma = 7.0
    pos = (rng.random((800, 2)) * [0.8, 0.3] + [-0.4, 0.0]).astype(np.float32)
    mass = np.full(800, 2.3, np.float32)
    b = O.bodies2(boxes=[((0.0, -0.05), (0.3, 0.05))])
    out = O.boundary2_f32(pos, mass, cfg, b)
    g = out["pos"].astype(np.float64)
    d2 = ((pos[None, :, :].astype(np.float64) - g[:, None, :]) ** 2).sum(axis=2)
    w = np.where(d2 < h * h, 4 / (math.pi * h ** 8) * np.maximum(h * h - d2, 0) ** 3, 0.0)
    rho = (w * mass[None, :]).sum(axis=1)
    assert (rho > 0).any()
    assert np.abs(out["density"] - np.minimum(rho, 3000.0)).max() <= 2e-5 * rho.max()


def _tank2():
    return O.bodies2(boxes=[((2.0, -1.0), (3.0, 0.5)), ((-1.0, 2.0), (0.5, 3.5)), ((5.0, 2.0), (0.5, 3.5))])


def test_reference_boundary_2d_present_and_contained(oracle_lib):
    """fluid.rs:1571-1630: ghosts exist and are deterministic; an 8 x 12
    column with ghosts stays finite, contained and calm for 300 steps."""
    pos = np.array([(ix * 0.25, iy * 0.25) for iy in range(12) for ix in range(8)], np.float32)
    mass = np.full(len(pos), 0.0625, np.float32)
    g1 = O.boundary2_f32(pos, mass, _cfg2(), _tank2())["pos"]
    g2 = O.boundary2_f32(pos, mass, _cfg2(), _tank2())["pos"]
    assert len(g1) > 0 and np.array_equal(g1.view(np.uint32), g2.view(np.uint32))
    p, v = pos, np.zeros_like(pos)
    peak = 0.0
    for _ in range(30):
        out = O.run2_f32(p, v, mass, _cfg2(), (0.0, -9.81), _tank2(), 1.0 / 60.0, 10)
        p, v = out["pos"], out["vel"]
        assert np.isfinite(p).all() and np.isfinite(v).all()
        peak = max(peak, float(np.linalg.norm(v, axis=1).max()))
    assert (p[:, 0] > -1).all() and (p[:, 0] < 5).all() and (p[:, 1] > -0.6).all()
    assert peak < 30.0

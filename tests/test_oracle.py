"""Pins of the CPU oracle (oracle/sph_oracle.cpp) — CPU only.

Each check is chosen so a plausible restatement mistake (a dropped term, a
wrong sign, a transposed index, a wrong neighbour order) fails: closed
forms, the reference's own known answers and behaviour tests
(crates/dynamics/src/fluid3.rs:1296-1605, fluid.rs:1326-1329), brute force
on small inputs, and invariants.
"""
import math

import numpy as np
import pytest

from phi_eng_amd import scenes
from phi_eng_amd.ffi import FluidConfig3Ffi


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def f32(x):
    return np.float32(x)


# ---------------------------------------------------------------- kernels
def test_poly6_zeros_signs_and_closed_form(O):
    """fluid3.rs:1296-1305 plus the closed form 315/(64 pi h^9)(h^2-r^2)^3."""
    h = 0.5
    peak = O.poly6(0.0, h)
    assert peak > 0
    assert O.poly6(h * h, h) == 0.0
    assert O.poly6(h * h * 4.0, h) == 0.0
    assert O.poly6(-1e-3, h) == 0.0
    assert O.poly6(h * h * 0.25, h) < peak
    for r2 in [0.0, 0.01, 0.1, 0.2]:
        exact = 315.0 / (64 * math.pi * h ** 9) * (h * h - r2) ** 3
        assert abs(O.poly6(r2, h) - exact) <= 2e-6 * exact
    # f64 twin agrees with the closed form to round-off
    assert abs(O.lib().oracle3_poly6_f64(0.01, 0.1) - 315 / (64 * math.pi * 0.1 ** 9) * (0.01 - 0.01) ** 3) < 1e-9


def test_spiky_grad_direction_and_zeros(O):
    """fluid3.rs:1306-1311 and -45/(pi h^6)(h-r)^2 r_vec/r."""
    h = 0.5
    g = O.spiky_grad([0.25, 0.0, 0.0], 0.25, h)
    assert g[0] < 0 and g[1] == 0 and g[2] == 0
    assert np.all(O.spiky_grad([h, 0.0, 0.0], h, h) == 0)
    assert np.all(O.spiky_grad([0.0, 0.0, 0.0], 0.0, h) == 0)
    rv = np.array([0.1, -0.2, 0.05])
    r = float(np.linalg.norm(rv))
    exact = -45 / (math.pi * h ** 6) * (h - r) ** 2 * rv / r
    np.testing.assert_allclose(O.spiky_grad(rv, r, h), exact, rtol=3e-6)


def test_eos_known_answers(O):
    """Linear EOS bit-level answer (fluid.rs:1326-1329) and Tait signs
    (fluid3.rs:1418-1425)."""
    k, rho0 = f32(100.0), f32(1.0)
    assert np.float32(O.pressure(1.2, 1.0, 100.0, 1.0)).view(np.uint32) == \
        (k * (f32(1.2) - rho0)).view(np.uint32)
    assert O.pressure(1.0, 1.0, 100.0, 7.0) == 0.0
    assert O.pressure(0.5, 1.0, 100.0, 7.0) == 0.0
    lin = O.pressure(1.2, 1.0, 100.0, 1.0)
    tait = O.pressure(1.2, 1.0, 100.0, 7.0)
    assert tait > lin > 0
    # Tait closed form k rho0 ((rho/rho0)^gamma - 1)
    assert abs(tait - 100 * (1.2 ** 7 - 1)) <= 1e-6 * tait
    assert abs(O.pressure(1010.0, 1000.0, 1.8e6, 7.0) - 1.8e9 * (1.01 ** 7 - 1)) <= 1e-6 * 1.8e9 * 0.072


# ---------------------------------------------------------------- grid
def test_cells_saturating_cast(O):
    """Rust `(x/h).floor() as i32`: saturating, NaN -> 0 (fluid3.rs:305-309)."""
    pos = np.array([[0.05, -0.05, 0.0], [1e30, -1e30, np.nan], [np.inf, -np.inf, 0.1],
                    [0.1, 0.2, 0.3], [-0.1, -0.2, -0.3]], dtype=np.float32)
    c = O.cells(pos, 0.1)
    q = np.floor(pos / np.float32(0.1)).astype(np.float64)  # f32 divide and floor, then widen
    q = np.where(np.isnan(q), 0.0, q)
    exp = np.clip(np.nan_to_num(q, posinf=2.0**40, neginf=-2.0**40), -2.0**31, 2.0**31 - 1)
    assert np.array_equal(c, exp.astype(np.int64).astype(np.int32))


def test_sorted_ids_is_stable_zmajor_sort(O):
    rng = np.random.default_rng(0)
    pos = rng.uniform(-0.3, 0.3, size=(500, 3)).astype(np.float32)
    pos[7] = [np.nan, 0, 0]
    pos[42] = [0, np.inf, 0]
    ids = O.sorted_ids(pos, 0.1)
    c = O.cells(pos, 0.1)
    fin = np.isfinite(pos).all(axis=1)
    key = np.lexsort((np.arange(len(pos)), c[:, 0], c[:, 1], c[:, 2], ~fin))
    assert np.array_equal(ids, key.astype(np.uint32))


# ---------------------------------------------------------------- density
def _brute(pos, mass, h, rho0_clamp):
    """O(N^2): neighbour counts with the reference's f32 cut-off test and the
    density as an f64 sum of f64 kernel values."""
    p = pos.astype(np.float32)
    d = p[None, :, :] - p[:, None, :]
    d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    nb = d2 < np.float32(h) * np.float32(h)
    r2 = np.sum((pos[None, :, :].astype(np.float64) - pos[:, None, :]) ** 2, axis=-1)
    w = 315.0 / (64 * math.pi * h ** 9) * np.clip(h * h - r2, 0, None) ** 3
    rho = np.clip((w * nb * mass[None, :]).sum(axis=1), 0, rho0_clamp)
    return nb.sum(axis=1), rho


def test_density_and_counts_brute_force(O):
    sc = scenes.small_dam(8, seed=21, jitter=0.05)
    pos = sc.positions
    m = O.masses(sc.config, len(pos))
    r = O.substep_f32(pos, np.zeros_like(pos), sc.config, sc.gravity, O.bodies_from_boxes(sc.boxes), 1e-4)
    cnt, rho = _brute(pos, m.astype(np.float64), 0.1, 3 * 1000.0)
    assert np.array_equal(r["count"], cnt)
    np.testing.assert_allclose(r["density"], rho, rtol=2e-6)


def test_lattice_density_closed_form(O):
    """Unjittered lattice, spacing h/2: an interior particle's density is
    m * 1.0097751668946855 / s^3 = rest density."""
    sc = scenes.small_dam(12, seed=0, jitter=0.0)
    pos = sc.positions
    r = O.substep_f32(pos, np.zeros_like(pos), sc.config, (0, 0, 0), O.bodies_from_boxes([]), 1e-4)
    c = np.array([6, 12, 6])
    idx = (c[2] * 24 + c[1]) * 12 + c[0]
    assert np.allclose(pos[idx], (c + 0.5) * 0.05)
    assert abs(r["density"][idx] - 1000.0) < 1e-3


# ---------------------------------------------------------------- force
def test_two_particle_force_closed_form(O):
    cfg = scenes.fluid_config(stiffness=2.0e4)
    cfg.viscosity = 0.0
    cfg.gamma = 1.0
    pos = np.array([[0.0, 0.0, 0.0], [0.06, 0.0, 0.0]], dtype=np.float32)
    r = O.substep_f32(pos, np.zeros_like(pos), cfg, (0, 0, 0), O.bodies_from_boxes([]), 1e-4)
    h, m = 0.1, float(cfg.phase_mass[0])
    w0 = 315 / (64 * math.pi * h ** 9) * h ** 6
    w1 = 315 / (64 * math.pi * h ** 9) * (h * h - 0.06 ** 2) ** 3
    rho = m * (w0 + w1)
    assert abs(r["density"][0] - rho) <= 2e-6 * rho
    p = max(0.0, 2.0e4 * (rho - 1000.0))
    grad = -45 / (math.pi * h ** 6) * (h - 0.06) ** 2  # along -x for particle 0
    f0 = -grad * (-1.0) * m * 2 * p / rho ** 2  # delta = x0 - x1 = -0.06 x
    a0 = f0 / rho
    assert abs(r["accel"][0, 0] - a0) <= 1e-5 * abs(a0) + 1e-6
    assert r["accel"][0, 0] == -r["accel"][1, 0]


def test_momentum_antisymmetry(O):
    """Symmetric pressure term: sum_i m_i F_i = 0 (no bodies, no gravity)."""
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 0.4, size=(400, 3)).astype(np.float32)
    cfg = scenes.fluid_config(stiffness=1e5)
    cfg.viscosity = 0.0
    r = O.substep_f32(pos, np.zeros_like(pos), cfg, (0, 0, 0), O.bodies_from_boxes([]), 1e-4)
    m = O.masses(cfg, len(pos)).astype(np.float64)
    f = r["accel"].astype(np.float64) * r["density"][:, None]
    tot = (m[:, None] * f).sum(axis=0)
    scale = (m[:, None] * np.abs(f)).sum(axis=0)
    assert np.all(np.abs(tot) <= 1e-5 * scale)


def test_f32_oracle_vs_fp64_twin(O):
    sc = scenes.small_dam(10, seed=23)
    pos = sc.positions
    b = O.bodies_from_boxes(sc.boxes)
    r32 = O.substep_f32(pos, np.zeros_like(pos), sc.config, sc.gravity, b, sc.dt)
    r64 = O.substep_f64(pos, np.zeros_like(pos), sc.config, sc.gravity, b, sc.dt)
    # The f32 path's own error against FP64: density ~1e-7; pressure and
    # acceleration are amplified by the Tait cancellation (rho/rho0)^7 - 1
    # (about 1/(7 * compression) ~ 1e3 here), so ~1e-4 of the force scale.
    ed = np.max(np.abs(r32["density"] - r64["density"]) / r64["density"])
    s = np.maximum(np.linalg.norm(r64["accel"], axis=1), 9.81 + r32["scale"])
    ea = np.max(np.abs(r32["accel"] - r64["accel"]).max(axis=1) / s)
    print(f"f32 oracle vs FP64: density {ed:.2e}, acceleration {ea:.2e} of scale")
    assert ed < 1e-5
    assert ea < 2e-3


def test_gauss_seidel_differs_from_jacobi_only_in_integrate(O):
    sc = scenes.small_dam(8, seed=24)
    pos = sc.positions
    rng = np.random.default_rng(1)
    vel = rng.normal(0, 0.3, size=pos.shape).astype(np.float32)
    b = O.bodies_from_boxes(sc.boxes)
    gs = O.substep_f32(pos, vel, sc.config, sc.gravity, b, sc.dt, jacobi=False)
    ja = O.substep_f32(pos, vel, sc.config, sc.gravity, b, sc.dt, jacobi=True)
    assert np.array_equal(gs["density"], ja["density"])
    assert np.array_equal(gs["accel"], ja["accel"])
    assert not np.array_equal(gs["vel"], ja["vel"])
    assert np.abs(gs["vel"] - ja["vel"]).max() < 0.1 * np.abs(vel).max()


# ------------------------------------------------- reference behaviour tests
def _tank_bodies(O):
    """fluid3.rs:1248-1262 tank: floor + four walls."""
    boxes = [((2.0, -1.0, 2.0), (3.0, 0.5, 3.0)),
             ((-1.0, 2.0, 2.0), (0.5, 3.5, 3.5)), ((5.0, 2.0, 2.0), (0.5, 3.5, 3.5)),
             ((2.0, 2.0, -1.0), (3.5, 3.5, 0.5)), ((2.0, 2.0, 5.0), (3.5, 3.5, 0.5))]
    return O.bodies_from_boxes(boxes)


def _lattice(origin, nx, ny, nz, dx):
    pts = [(origin[0] + ix * dx, origin[1] + iy * dx, origin[2] + iz * dx)
           for iz in range(nz) for iy in range(ny) for ix in range(nx)]
    return np.array(pts, dtype=np.float32)


def test_reference_dam_break_contained(O):
    """fluid3.rs:1313-1347 with the default config (Gauss-Seidel oracle)."""
    cfg = FluidConfig3Ffi.default()
    pos = _lattice((0, 0, 0), 4, 6, 4, 0.25)
    vel = np.zeros_like(pos)
    b = _tank_bodies(O)
    peak, max_x = 0.0, 0.0
    for frame in range(400):
        pos, vel, _ = O.run_f32(pos, vel, cfg, (0, -9.81, 0), b, 1 / 60, 1)
        peak = max(peak, float(np.linalg.norm(vel, axis=1).max()))
        max_x = max(max_x, float(pos[:, 0].max()))
    assert np.isfinite(pos).all() and np.isfinite(vel).all()
    assert max_x > 1.5
    assert (pos[:, 0] > -1).all() and (pos[:, 0] < 5).all() and (pos[:, 2] > -1).all() \
        and (pos[:, 2] < 5).all() and (pos[:, 1] > -0.6).all()
    assert peak < 30
    assert len(pos) == 96


def test_reference_determinism(O):
    """fluid3.rs:1398-1416: two runs are bit-identical."""
    cfg = FluidConfig3Ffi.default()
    pos = _lattice((0, 0, 0), 4, 5, 4, 0.25)
    b = _tank_bodies(O)
    a = O.run_f32(pos, np.zeros_like(pos), cfg, (0, -9.81, 0), b, 1 / 60, 80)
    c = O.run_f32(pos, np.zeros_like(pos), cfg, (0, -9.81, 0), b, 1 / 60, 80)
    assert np.array_equal(a[0].view(np.uint32), c[0].view(np.uint32))


def test_reference_cohesion_attracts(O):
    """fluid3.rs:1457-1471."""
    cfg = FluidConfig3Ffi.default()
    cfg.stiffness = 0.0
    cfg.viscosity = 0.0
    cfg.surface_tension = 1000.0
    pos = np.array([[0, 0, 0], [0.4, 0, 0]], dtype=np.float32)
    p, v, _ = O.run_f32(pos, np.zeros_like(pos), cfg, (0, 0, 0), O.bodies_from_boxes([]), 1 / 60, 1)
    assert v[0, 0] > 0 and v[1, 0] < 0


def test_reference_cfl_substeps(O):
    """fluid3.rs:1473-1487."""
    cfg = FluidConfig3Ffi.default()
    vel = np.zeros((2, 3), dtype=np.float32)
    assert O.cfl_substeps(vel, cfg, 1 / 60) == 1
    cfg.gamma = 7.0
    cfg.stiffness = 200.0
    cfg.cfl = 0.4
    cfg.max_substeps = 8
    vel[0, 0] = 20.0
    assert 2 <= O.cfl_substeps(vel, cfg, 1 / 60) <= 8


def test_reference_hydrostatic_band(O):
    """fluid3.rs:1560-1605: still column, p_bottom in [0.25, 3.5] rho g h."""
    cfg = FluidConfig3Ffi.default()
    pos = _lattice((0, -0.375, 0), 8, 10, 8, 0.25)
    b = _tank_bodies(O)
    pos, vel, _ = O.run_f32(pos, np.zeros_like(pos), cfg, (0, -9.81, 0), b, 1 / 60, 900)
    r = O.substep_f32(pos, vel, cfg, (0, -9.81, 0), b, 1e-9, jacobi=False)
    ys = np.sort(pos[(pos[:, 1] > -1) & (pos[:, 1] < 6), 1])
    n, tenth = len(ys), len(ys) // 10
    bottom, surface = ys[:tenth].mean(), ys[n - tenth:].mean()
    lo, hi = ys[tenth], ys[n - tenth]
    p = r["pressure"]
    p_bottom = p[pos[:, 1] <= lo].mean()
    p_top = p[pos[:, 1] >= hi].mean()
    expected = 1.0 * 9.81 * (surface - bottom)
    assert p_bottom > p_top
    assert 0.25 * expected < p_bottom < 3.5 * expected


def test_boundary_shapes_push_out(O):
    """collide_particle_body (fluid3.rs:471-585): a particle inside a sphere,
    a box (least-penetration axis, always +axis) and a capsule is pushed to
    the surface plus radius and loses its approaching normal velocity."""
    cfg = scenes.fluid_config(stiffness=0.0)
    cfg.viscosity = 0.0
    cfg.particle_radius = 0.01
    # sphere radius 0.5 at origin; particle at 0.3 on +x moving -x
    b = O.bodies_from_boxes([], spheres=[((0, 0, 0), 0.5)])
    r = O.substep_f32(np.array([[0.3, 0, 0]], np.float32), np.array([[-1, 0, 0]], np.float32), cfg, (0, 0, 0),
                      b, 1e-6)
    assert abs(r["pos"][0, 0] - 0.51) < 1e-5 and r["vel"][0, 0] == 0
    # box: centre inside, nearest face -x, but ejection is along +x (quirk)
    b = O.bodies_from_boxes([((0, 0, 0), (0.5, 0.5, 0.5))])
    r = O.substep_f32(np.array([[-0.45, 0, 0]], np.float32), np.zeros((1, 3), np.float32), cfg, (0, 0, 0),
                      b, 1e-6)
    assert r["pos"][0, 0] > -0.45
    # capsule along y: radial push
    b = O.bodies_from_boxes([], capsules=[((0, 0, 0), 0.2, 0.5)])
    r = O.substep_f32(np.array([[0.0, 0.1, 0.15]], np.float32), np.zeros((1, 3), np.float32), cfg, (0, 0, 0),
                      b, 1e-6)
    assert abs(np.hypot(r["pos"][0, 0], r["pos"][0, 2]) - 0.21) < 1e-5


# ---------------------------------------------------------------- golden
def test_golden_fixtures_reproduce(O):
    """Fixtures under tests/golden were written by make_golden.py from this
    oracle; a silent change to the oracle's arithmetic fails here."""
    import os
    from tests.golden import make_golden
    path = os.path.join(os.path.dirname(__file__), "golden")
    files = sorted(f for f in os.listdir(path) if f.endswith(".npz"))
    assert files, "no golden fixtures committed"
    for f in files:
        g = np.load(os.path.join(path, f))
        out = make_golden.recompute(g)
        for k in ("density", "pressure", "accel", "count", "pos_out", "vel_out"):
            assert np.array_equal(out[k], g[k]), (f, k)


def test_openmp_jacobi_equals_serial(oracle_lib):
    """The all-cores baseline (oracle3_run_f32_par) is the one-thread Jacobi
    oracle on OpenMP threads: equal bit for bit, with AV and cohesion on."""
    from phi_eng_amd import scenes
    O = oracle_lib
    sc = scenes.small_dam(10, seed=41)
    sc.config.artificial_viscosity = 0.3
    sc.config.surface_tension = 0.1
    rng = np.random.default_rng(41)
    vel = rng.uniform(-0.5, 0.5, sc.positions.shape).astype(np.float32)
    b = O.bodies_from_boxes(sc.boxes)
    a = O.run_f32(sc.positions, vel, sc.config, sc.gravity, b, 5e-4, 4, jacobi=True)
    p = O.run_f32_par(sc.positions, vel, sc.config, sc.gravity, b, 5e-4, 4, threads=4)
    assert np.array_equal(a[0].view(np.uint32), p[0].view(np.uint32))
    assert np.array_equal(a[1].view(np.uint32), p[1].view(np.uint32))

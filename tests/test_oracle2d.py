"""Pins of the 2D CPU oracle (oracle/sph2d_oracle.cpp, a restatement of
crates/dynamics/src/fluid.rs): the reference's own 2D tests restated, the
kernels' closed forms and normalisation, brute-force neighbour counts and
densities, the lattice closed form, and momentum antisymmetry."""
import math

import numpy as np
import pytest

from oracle import oracle as O

f32 = np.float32


def tank():
    """fluid.rs:1151-1171: floor and two walls."""
    return O.bodies2(boxes=[((2.0, -1.0), (3.0, 0.5)), ((-1.0, 2.0), (0.5, 3.5)), ((5.0, 2.0), (0.5, 3.5))])


def lattice(x0, y0, nx, ny, dx):
    """fluid.rs:1173-1182 (row-major from (x0, y0))."""
    return np.array([(x0 + ix * dx, y0 + iy * dx) for iy in range(ny) for ix in range(nx)], dtype=np.float32)


def test_kernels_sane_reference():
    """fluid.rs:1209-1223 restated."""
    h = 0.5
    peak = O.poly6_2(0.0, h)
    assert peak > 0.0
    assert O.poly6_2(h * h, h) == 0.0
    assert O.poly6_2(h * h * 4.0, h) == 0.0
    assert O.poly6_2(h * h * 0.25, h) < peak
    g = O.spiky_grad2(0.25, 0.0, 0.25, h)
    assert g[0] < 0.0 and g[1] == 0.0
    assert (O.spiky_grad2(h, 0.0, h, h) == 0.0).all()
    assert (O.spiky_grad2(0.0, 0.0, 0.0, h) == 0.0).all()


def test_kernel_closed_forms_and_normalisation():
    h = 0.1
    h32 = float(f32(h))
    h2_32 = float(f32(h32) * f32(h32))  # the kernel's h * h in f32
    for r in (0.0, 0.013, 0.05, 0.0999):
        r2 = float(f32(r * r))
        w = 4.0 / (math.pi * h32 ** 8) * (h2_32 - r2) ** 3  # exact t = h2 - r2 (Sterbenz)
        assert abs(O.poly6_2(r2, h) - w) <= 2e-6 * w
        if r > 0:
            r32 = float(f32(r))
            g = 30.0 / (math.pi * h32 ** 5) * (h32 - r32) ** 2  # exact h - r (Sterbenz)
            out = O.spiky_grad2(r32, 0.0, r32, h)
            assert abs(-out[0] - g) <= 2e-6 * g and out[1] == 0.0
    # the poly6 kernel integrates to 1 over the disk of radius h
    k = 801
    xs = (np.arange(k) - (k - 1) / 2) * (2 * h / (k - 1))
    X, Y = np.meshgrid(xs, xs)
    r2 = (X * X + Y * Y).ravel()
    total = sum(O.poly6_2(float(v), h) for v in r2[r2 < h * h]) * (xs[1] - xs[0]) ** 2
    assert abs(total - 1.0) < 2e-3


def test_tait_eos_known_answers():
    """fluid.rs:1311-1331 restated, including the bit-level linear EOS."""
    k, rho0 = 100.0, 1.0
    assert O.pressure2(rho0, rho0, k, 7.0) == 0.0
    assert O.pressure2(0.5, rho0, k, 7.0) == 0.0
    lin = O.pressure2(1.2, rho0, k, 1.0)
    tait = O.pressure2(1.2, rho0, k, 7.0)
    assert lin > 0.0 and tait > lin
    assert f32(lin).view(np.uint32) == (f32(k) * (f32(1.2) - f32(rho0))).view(np.uint32)
    # Tait closed form at rho = 1.2 (f64 power rounded to f32)
    want = max(k * rho0 * (1.2 ** 7 - 1.0), 0.0)
    assert abs(tait - want) <= 1e-6 * want


def _brute(pos, h, mass, rho_cap):
    n = len(pos)
    d = pos[None, :, :] - pos[:, None, :]  # f32 pj - pi
    d2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]
    nb = d2 < f32(h) * f32(h)
    cnt = nb.sum(axis=1)
    r2 = d2.astype(np.float64)
    w = np.where(nb, 4.0 / (math.pi * h ** 8) * np.maximum(h * h - r2, 0.0) ** 3, 0.0)
    rho = np.clip((w * mass[None, :].astype(np.float64)).sum(axis=1), 0.0, rho_cap)
    return cnt, rho


def test_brute_force_counts_and_density():
    rng = np.random.default_rng(3)
    cfg = O.Oracle2Cfg.default()
    cfg.smoothing_length = 0.1
    cfg.rest_density = 1000.0
    for k in range(4):
        cfg.phase_rest_density[k] = 1000.0
    pos = (rng.random((700, 2)) * 0.8 - 0.13).astype(np.float32)  # also negative cells
    mass = np.full(700, 2.3, np.float32)
    out = O.substep2_f32(pos, np.zeros_like(pos), mass, cfg, (0.0, 0.0), O.bodies2(), 1e-4)
    cnt, rho = _brute(pos, 0.1, mass, 3000.0)
    assert np.array_equal(out["count"], cnt)
    assert np.abs(out["density"] - rho).max() <= 2e-6 * rho.max()


def test_lattice_closed_form():
    """Interior density of a square lattice at s = h/2 = m * sum_j W(|x_j|)."""
    h, s, m = 0.1, 0.05, 1.7
    cfg = O.Oracle2Cfg.default()
    cfg.smoothing_length = h
    cfg.rest_density = 1e4
    g = np.array([(ix * s, iy * s) for iy in range(11) for ix in range(11)], np.float32)
    out = O.substep2_f32(g, np.zeros_like(g), np.full(len(g), m, np.float32), cfg, (0.0, 0.0), O.bodies2(), 1e-4)
    lat = [(a * s, b * s) for a in range(-2, 3) for b in range(-2, 3)]
    want = m * sum(4.0 / (math.pi * h ** 8) * max(h * h - (x * x + y * y), 0.0) ** 3 for x, y in lat)
    centre = 5 * 11 + 5
    assert abs(out["density"][centre] - want) <= 2e-6 * want


def test_momentum_antisymmetry():
    """Pressure forces are pairwise antisymmetric: sum_i rho_i m_i (a_i - g) = 0."""
    rng = np.random.default_rng(5)
    cfg = O.Oracle2Cfg.default()
    cfg.smoothing_length = 0.1
    cfg.rest_density = 1000.0
    cfg.gamma = 7.0
    cfg.stiffness = 50.0
    cfg.viscosity = 0.0
    for k in range(4):
        cfg.phase_rest_density[k] = 1000.0
    pos = (rng.random((400, 2)) * 0.5).astype(np.float32)
    mass = np.full(400, 3.0, np.float32)
    out = O.substep2_f32(pos, np.zeros_like(pos), mass, cfg, (0.0, -9.81), O.bodies2(), 1e-4)
    f = (out["accel"].astype(np.float64) - [0.0, -9.81]) * out["density"][:, None] * mass[:, None]
    scale = np.abs(f).sum()
    assert scale > 0
    assert np.abs(f.sum(axis=0)).max() <= 1e-5 * scale


def _default_world(pos):
    cfg = O.Oracle2Cfg.default()
    return cfg, np.full(len(pos), 0.0625, np.float32)  # FluidConfig2::default phase_mass


def test_dam_break_contained():
    """fluid.rs:1225-1256 restated: 8 x 12 column, 600 steps at 60 Hz."""
    pos = lattice(0.0, 0.0, 8, 12, 0.25)
    vel = np.zeros_like(pos)
    cfg, mass = _default_world(pos)
    b = tank()
    peak = max_x = 0.0
    for frame in range(600):
        out = O.run2_f32(pos, vel, mass, cfg, (0.0, -9.81), b, 1.0 / 60.0, 1)
        pos, vel = out["pos"], out["vel"]
        assert np.isfinite(pos).all() and np.isfinite(vel).all()
        peak = max(peak, float(np.sqrt((vel.astype(np.float64) ** 2).sum(axis=1)).max()))
        max_x = max(max_x, float(pos[:, 0].max()))
    assert max_x > 2.5, f"dam never collapsed, max_x={max_x}"
    assert (pos[:, 0] > -1.0).all() and (pos[:, 0] < 5.0).all() and (pos[:, 1] > -0.6).all()
    assert peak < 30.0
    assert len(pos) == 96


def test_determinism():
    """fluid.rs:1292-1309: the same scene run twice is bit-identical."""
    pos = lattice(0.0, 0.0, 8, 12, 0.25)
    cfg, mass = _default_world(pos)
    a = O.run2_f32(pos, np.zeros_like(pos), mass, cfg, (0.0, -9.81), tank(), 1.0 / 60.0, 200)
    b = O.run2_f32(pos, np.zeros_like(pos), mass, cfg, (0.0, -9.81), tank(), 1.0 / 60.0, 200)
    assert np.array_equal(a["pos"].view(np.uint32), b["pos"].view(np.uint32))
    assert np.array_equal(a["vel"].view(np.uint32), b["vel"].view(np.uint32))


def test_hydrostatic_band():
    """fluid.rs:1676-1719 restated: still column, pressure rises with depth
    and the bottom decile's pressure is within [0.3, 3] x rho g h."""
    pos = lattice(0.0, 0.0, 18, 12, 0.25)
    cfg, mass = _default_world(pos)
    out = O.run2_f32(pos, np.zeros_like(pos), mass, cfg, (0.0, -9.81), tank(), 1.0 / 60.0, 1200)
    ys = np.sort(out["pos"][:, 1])
    n, tenth = len(ys), len(ys) // 10
    bottom, surface = ys[:tenth].mean(), ys[n - tenth:].mean()
    lo, hi = ys[tenth], ys[n - tenth]
    y, p = out["pos"][:, 1], out["pressure"]
    p_bottom, p_top = p[y <= lo].mean(), p[y >= hi].mean()
    expected = 1.0 * 9.81 * (surface - bottom)
    assert p_bottom > p_top
    assert 0.3 * expected < p_bottom < 3.0 * expected


def test_jacobi_and_gauss_seidel_share_density():
    """The integrate mode changes only the integrate: the substep's density,
    pressure and acceleration are identical."""
    pos = lattice(0.0, 0.0, 8, 12, 0.25)
    cfg, mass = _default_world(pos)
    vel = np.random.default_rng(1).normal(0, 0.5, pos.shape).astype(np.float32)
    a = O.substep2_f32(pos, vel, mass, cfg, (0.0, -9.81), tank(), 1.0 / 60.0, jacobi=False)
    b = O.substep2_f32(pos, vel, mass, cfg, (0.0, -9.81), tank(), 1.0 / 60.0, jacobi=True)
    for k in ("density", "pressure", "count"):
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["vel"], b["vel"])  # the integrate differs


def test_config1_dam_stays_physical():
    """The 2D dam-break workload (4,096 particles) on the oracle: 300 steps
    keep the fluid near rest density and inside the tank."""
    from phi_eng_amd import scenes
    sc = scenes.dam2d_config1()
    cfg, mass, b = O.scene2_inputs(sc)
    out = O.run2_f32(sc.positions, np.zeros_like(sc.positions), mass, cfg, sc.gravity, b, sc.dt, 300)
    d = out["density"]
    assert np.isfinite(out["pos"]).all()
    assert 980.0 < np.median(d) < 1080.0 and d.max() < 1300.0
    x, y = out["pos"][:, 0], out["pos"][:, 1]
    assert (x > 0.0).all() and (x < 12.8).all() and (y > 0.0).all()
    assert x.max() > 4.0  # the column has collapsed rightward

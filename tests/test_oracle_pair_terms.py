"""Pins of the oracle's optional pair terms — CPU only.

Artificial viscosity (fluid3.rs:254-269, fluid.rs:302-317), cohesion
(fluid3.rs:272-278 and 1062-1070, fluid.rs:320-327) and XSPH smoothing
(fluid3.rs:1112-1123, fluid.rs:1021-1040) in both oracles
(oracle/sph_oracle.cpp, oracle/sph2d_oracle.cpp):

- the reference's own sign tests restated (fluid3.rs:1418-1455,
  fluid.rs:1333-1360);
- closed forms of a whole substep of two particles of different masses,
  one term switched on at a time (pressure held at zero by a rest density
  far above the pair's density), so a wrong sign, a dropped factor or
  m_i for m_j fails;
- invariants: the pair's momentum is conserved by each term, separating
  pairs feel no viscosity, equal velocities feel no smoothing, and the
  reference's in-place (Gauss-Seidel) XSPH reads the first particle's new
  velocity.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from phi_eng_amd.ffi import FluidConfig3Ffi

# config values as the f32 fields hold them, so the FP64 closed forms and
# the FP64 twin start from the same numbers
H = float(np.float32(0.1))
M0, M1 = float(np.float32(0.12)), float(np.float32(0.2))  # unequal masses: m_i / m_j mix-ups fail
POS3 = np.array([[0.0, 0.0, 0.0], [0.06, 0.01, -0.02]], np.float32)
VEL3 = np.array([[1.0, 0.5, 0.0], [-1.0, 0.0, 0.3]], np.float32)
PH = np.array([0, 1], np.uint8)


@pytest.fixture(scope="module", autouse=True)
def _built(oracle_lib):
    return oracle_lib


# ---------------------------------------------------------------- signs
def test_reference_av_and_cohesion_signs_3d():
    """fluid3.rs:1426-1453 restated."""
    h = 0.5
    assert O.cohesion(0.1, h) > O.cohesion(0.4, h)
    assert O.cohesion(h, h) == 0.0
    delta = [1.0, 0.0, 0.0]
    assert O.artificial_pressure(delta, [-2.0, 0.0, 0.0], 1.0, 0.5, 1.0, 10.0, 1.0) > 0.0
    assert O.artificial_pressure(delta, [2.0, 0.0, 0.0], 1.0, 0.5, 1.0, 10.0, 1.0) == 0.0


def test_reference_av_and_cohesion_signs_2d():
    """fluid.rs:1333-1360 restated (test_fluid_artificial_pressure_sign,
    test_fluid_cohesion_kernel)."""
    d = (1.0, 0.0)
    assert O.artificial_pressure2(d, (-2.0, 0.0), 1.0, 0.5, 1.0, 10.0, 1.0) > 0.0
    assert O.artificial_pressure2(d, (2.0, 0.0), 1.0, 0.5, 1.0, 10.0, 1.0) == 0.0
    assert O.artificial_pressure2(d, (-2.0, 0.0), 1.0, 0.5, 1.0, 10.0, 0.0) == 0.0
    h = 0.5
    assert O.cohesion2(0.1, h) > O.cohesion2(0.4, h)
    assert O.cohesion2(h, h) == 0.0
    assert O.cohesion2(0.9, h) == 0.0


def test_av_scalar_values():
    """The Monaghan term's value on a hand-worked case: vd = -2, mu =
    0.5 * -2 / (1 + 0.0025) and -alpha c mu / rho_bar with alpha 1, c 10,
    rho_bar 1 (3D and 2D)."""
    mu = 0.5 * -2.0 / (1.0 + 0.01 * 0.25)
    want = -1.0 * 10.0 * mu / 1.0
    got3 = O.artificial_pressure([1.0, 0.0, 0.0], [-2.0, 0.0, 0.0], 1.0, 0.5, 1.0, 10.0, 1.0)
    got2 = O.artificial_pressure2((1.0, 0.0), (-2.0, 0.0), 1.0, 0.5, 1.0, 10.0, 1.0)
    assert abs(got3 - want) <= 2e-6 * abs(want)
    assert abs(got2 - want) <= 2e-6 * abs(want)
    # rho_bar below EPS is lifted to EPS (fluid3.rs:268)
    tiny = O.artificial_pressure([1.0, 0.0, 0.0], [-2.0, 0.0, 0.0], 1.0, 0.5, 0.0, 10.0, 1.0)
    assert abs(tiny - want / 1e-6) <= 2e-6 * abs(want / 1e-6)


# ------------------------------------------------------- 3D closed forms
def _cfg3(k=1.0e4, av=0.0, kappa=0.0, nu=0.0):
    c = FluidConfig3Ffi.default()
    c.smoothing_length = H
    c.rest_density = 1.0e6
    c.particle_mass = M0
    c.phase_count = 2
    c.phase_rest_density[0] = c.phase_rest_density[1] = 1.0e6
    c.phase_mass[0], c.phase_mass[1] = M0, M1
    c.stiffness = k
    c.gamma = 7.0
    c.artificial_viscosity = av
    c.surface_tension = kappa
    c.viscosity = nu
    c.particle_radius = 0.0
    return c


def _w3(r2):
    return 315.0 / (64.0 * math.pi * H ** 9) * (H * H - r2) ** 3 if r2 < H * H else 0.0


def _pair3():
    x = POS3.astype(np.float64)
    d = x[0] - x[1]  # delta of particle 0 (x_i - x_j)
    r2 = float(d @ d)
    r = math.sqrt(r2)
    w0, w1 = _w3(0.0), _w3(r2)
    rho = np.array([M0 * w0 + M1 * w1, M1 * w0 + M0 * w1])
    grad0 = d / r * (-45.0 / (math.pi * H ** 6)) * (H - r) ** 2
    return d, r2, r, w1, rho, grad0


def _run3(cfg, jacobi=True):
    b = O.bodies_from_boxes([])
    return O.substep_f32(POS3, VEL3, cfg, (0, 0, 0), b, 1e-4, jacobi=jacobi, phases=PH)


def test_pair_densities_and_zero_pressure_3d():
    _, _, _, _, rho, _ = _pair3()
    r = _run3(_cfg3())
    np.testing.assert_allclose(r["density"], rho, rtol=2e-6)
    assert np.all(r["pressure"] == 0.0)
    np.testing.assert_array_equal(r["accel"], 0.0)


def test_av_two_particle_closed_form_3d():
    """An approaching pair, AV alone: a_i = -grad W_ij m_j Pi / rho_i with
    Pi = -alpha c mu / rho_bar, mu = h vd / (r^2 + 0.01 h^2), c = sqrt(k)."""
    k, alpha = 1.0e4, 0.5
    d, r2, r, _, rho, grad0 = _pair3()
    v = VEL3.astype(np.float64)
    vd = float((v[0] - v[1]) @ d)
    assert vd < 0.0
    mu = H * vd / (r2 + 0.01 * H * H)
    pi_ = -alpha * math.sqrt(k) * mu / (0.5 * (rho[0] + rho[1]))
    assert pi_ > 0.0
    a0 = -grad0 * M1 * pi_ / rho[0]
    a1 = grad0 * M0 * pi_ / rho[1]
    got = _run3(_cfg3(k=k, av=alpha))
    np.testing.assert_allclose(got["accel"][0], a0, rtol=1e-5, atol=1e-6 * np.abs(a0).max())
    np.testing.assert_allclose(got["accel"][1], a1, rtol=1e-5, atol=1e-6 * np.abs(a1).max())
    # repulsive: particle 0 is pushed away from particle 1
    assert float(got["accel"][0].astype(np.float64) @ d) > 0.0
    # momentum: m0 rho0 a0 + m1 rho1 a1 = m0 F0 + m1 F1 = 0
    f = got["accel"].astype(np.float64) * got["density"][:, None]
    tot = M0 * f[0] + M1 * f[1]
    assert np.abs(tot).max() <= 1e-5 * (M0 * np.abs(f[0])).max()
    # the FP64 twin follows the same closed form to round-off
    t = O.substep_f64(POS3, VEL3, _cfg3(k=k, av=alpha), (0, 0, 0), O.bodies_from_boxes([]), 1e-4, phases=PH)
    np.testing.assert_allclose(t["accel"][0], a0, rtol=1e-10)


def test_av_separating_pair_feels_nothing_3d():
    vel = -VEL3
    b = O.bodies_from_boxes([])
    r = O.substep_f32(POS3, vel, _cfg3(k=1.0e4, av=0.5), (0, 0, 0), b, 1e-4, phases=PH)
    np.testing.assert_array_equal(r["accel"], 0.0)


def test_cohesion_two_particle_closed_form_3d():
    """a_i = -(delta / r) kappa m_i m_j (1 - r / h) / rho_i (fluid3.rs:1062-1070)."""
    kappa = 50.0
    d, _, r, _, rho, _ = _pair3()
    c = 1.0 - r / H
    a0 = -(d / r) * kappa * M0 * M1 * c / rho[0]
    a1 = (d / r) * kappa * M1 * M0 * c / rho[1]
    got = _run3(_cfg3(kappa=kappa))
    np.testing.assert_allclose(got["accel"][0], a0, rtol=1e-5, atol=1e-6 * np.abs(a0).max())
    np.testing.assert_allclose(got["accel"][1], a1, rtol=1e-5, atol=1e-6 * np.abs(a1).max())
    # attractive
    assert float(got["accel"][0].astype(np.float64) @ d) < 0.0
    # beyond h nothing
    far = POS3.copy()
    far[1] = [0.11, 0.0, 0.0]
    r2 = O.substep_f32(far, VEL3, _cfg3(kappa=kappa), (0, 0, 0), O.bodies_from_boxes([]), 1e-4, phases=PH)
    np.testing.assert_array_equal(r2["accel"], 0.0)


def test_xsph_two_particle_closed_form_3d():
    """Jacobi: v_i' = v_i + nu (v_j - v_i) (m_j / rho_bar) W(r^2); the
    reference's in-place sweep (fluid3.rs:1092-1139) gives particle 1 the
    smoothing against particle 0's new velocity."""
    nu = 0.3
    _, r2, _, w1, rho, _ = _pair3()
    rb = 0.5 * (rho[0] + rho[1])
    v = VEL3.astype(np.float64)
    v0 = v[0] + nu * (v[1] - v[0]) * (M1 / rb) * w1
    v1 = v[1] + nu * (v[0] - v[1]) * (M0 / rb) * w1
    got = _run3(_cfg3(nu=nu), jacobi=True)
    np.testing.assert_allclose(got["vel"][0], v0, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(got["vel"][1], v1, rtol=1e-6, atol=1e-7)
    # the pair's momentum is conserved by the smoothing
    p0 = M0 * v[0] + M1 * v[1]
    p1 = M0 * got["vel"][0].astype(np.float64) + M1 * got["vel"][1].astype(np.float64)
    np.testing.assert_allclose(p1, p0, rtol=0, atol=1e-6)
    # relative velocity shrinks
    assert np.linalg.norm(got["vel"][0] - got["vel"][1]) < np.linalg.norm(v[0] - v[1])
    # Gauss-Seidel: particle 1 smooths against particle 0's new velocity
    # and new position x0 + v0' dt (both overwritten in place first)
    gs = _run3(_cfg3(nu=nu), jacobi=False)
    x0n = POS3[0].astype(np.float64) + v0 * 1e-4
    dn = x0n - POS3[1].astype(np.float64)
    v1_gs = v[1] + nu * (v0 - v[1]) * (M0 / rb) * _w3(float(dn @ dn))
    np.testing.assert_allclose(gs["vel"][0], v0, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(gs["vel"][1], v1_gs, rtol=1e-6, atol=1e-7)
    assert not np.allclose(v1_gs, v1, rtol=0, atol=1e-9)
    # equal velocities: nothing to smooth
    same = np.repeat(VEL3[:1], 2, axis=0)
    r = O.substep_f32(POS3, same, _cfg3(nu=nu), (0, 0, 0), O.bodies_from_boxes([]), 1e-4, phases=PH)
    np.testing.assert_array_equal(r["vel"], same)


def _cloud():
    rng = np.random.default_rng(11)
    pos = rng.uniform(0.0, 0.3, size=(300, 3)).astype(np.float32)
    vel = rng.normal(0.0, 0.5, size=pos.shape).astype(np.float32)
    ph = (rng.uniform(size=300) < 0.5).astype(np.uint8)
    return pos, vel, ph


def test_all_terms_together_3d_f64_twin():
    """AV, cohesion and XSPH together on a random cloud of two masses: the
    f32 oracle stays within 1e-4 of the acceleration scale of its FP64
    twin, and velocities within 1e-5."""
    pos, vel, ph = _cloud()
    cfg = _cfg3(k=1.0e4, av=0.4, kappa=20.0, nu=0.3)
    b = O.bodies_from_boxes([])
    r32 = O.substep_f32(pos, vel, cfg, (0, 0, 0), b, 1e-4, phases=ph)
    r64 = O.substep_f64(pos, vel, cfg, (0, 0, 0), b, 1e-4, phases=ph)
    s = np.maximum(np.linalg.norm(r64["accel"], axis=1), r32["scale"])
    assert np.max(np.abs(r32["accel"] - r64["accel"]).max(axis=1) / np.maximum(s, 1e-30)) < 1e-4
    np.testing.assert_allclose(r32["vel"], r64["vel"], rtol=0, atol=1e-5)


def test_pair_term_invariants_on_a_cloud_3d():
    """The AV term is symmetric in (i, j) times grad W (antisymmetric), so
    with pressure held at zero sum_i m_i F_i = 0; the cohesion term is
    kappa m_i m_j times an antisymmetric direction, so sum_i F_i = 0."""
    pos, vel, ph = _cloud()
    m = np.where(ph == 0, M0, M1)
    b = O.bodies_from_boxes([])
    av = O.substep_f64(pos, vel, _cfg3(k=1.0e4, av=0.4), (0, 0, 0), b, 1e-4, phases=ph)
    f = av["accel"] * av["density"][:, None]
    assert np.abs(f).max() > 0.0
    assert np.abs((m[:, None] * f).sum(axis=0)).max() <= 1e-9 * (m[:, None] * np.abs(f)).sum()
    co = O.substep_f64(pos, vel, _cfg3(kappa=20.0), (0, 0, 0), b, 1e-4, phases=ph)
    f = co["accel"] * co["density"][:, None]
    assert np.abs(f).max() > 0.0
    assert np.abs(f.sum(axis=0)).max() <= 1e-9 * np.abs(f).sum()


# ------------------------------------------------------- 2D closed forms
POS2 = np.array([[0.0, 0.0], [0.06, 0.02]], np.float32)
VEL2 = np.array([[1.0, 0.5], [-1.0, 0.2]], np.float32)
MASS2 = np.array([M0, M1], np.float32)


def _cfg2(k=1.0e4, av=0.0, kappa=0.0, nu=0.0):
    c = O.Oracle2Cfg.default()
    c.smoothing_length = H
    c.rest_density = 1.0e6
    for q in range(4):
        c.phase_rest_density[q] = 1.0e6
    c.stiffness = k
    c.gamma = 7.0
    c.artificial_viscosity = av
    c.surface_tension = kappa
    c.viscosity = nu
    c.particle_radius = 0.0
    c.particle_mass = M0
    return c


def _w2(r2):
    return 4.0 / (math.pi * H ** 8) * (H * H - r2) ** 3 if r2 < H * H else 0.0


def _pair2():
    x = POS2.astype(np.float64)
    d = x[0] - x[1]
    r2 = float(d @ d)
    r = math.sqrt(r2)
    w0, w1 = _w2(0.0), _w2(r2)
    rho = np.array([M0 * w0 + M1 * w1, M1 * w0 + M0 * w1])
    grad0 = d / r * (-30.0 / (math.pi * H ** 5)) * (H - r) ** 2
    return d, r2, r, w1, rho, grad0


def _run2(cfg, jacobi=True, vel=VEL2, pos=POS2):
    return O.substep2_f32(pos, vel, MASS2, cfg, (0.0, 0.0), O.bodies2(), 1e-4, jacobi=jacobi)


def test_av_two_particle_closed_form_2d():
    k, alpha = 1.0e4, 0.5
    d, r2, r, _, rho, grad0 = _pair2()
    got = _run2(_cfg2(k=k, av=alpha))
    np.testing.assert_allclose(got["density"], rho, rtol=2e-6)
    assert np.all(got["pressure"] == 0.0)
    v = VEL2.astype(np.float64)
    vd = float((v[0] - v[1]) @ d)
    assert vd < 0.0
    mu = H * vd / (r2 + 0.01 * H * H)
    pi_ = -alpha * math.sqrt(k) * mu / (0.5 * (rho[0] + rho[1]))
    a0 = -grad0 * M1 * pi_ / rho[0]
    a1 = grad0 * M0 * pi_ / rho[1]
    np.testing.assert_allclose(got["accel"][0], a0, rtol=1e-5, atol=1e-6 * np.abs(a0).max())
    np.testing.assert_allclose(got["accel"][1], a1, rtol=1e-5, atol=1e-6 * np.abs(a1).max())
    assert float(got["accel"][0].astype(np.float64) @ d) > 0.0
    sep = _run2(_cfg2(k=k, av=alpha), vel=-VEL2)
    np.testing.assert_array_equal(sep["accel"], 0.0)


def test_cohesion_two_particle_closed_form_2d():
    kappa = 50.0
    d, _, r, _, rho, _ = _pair2()
    c = 1.0 - r / H
    a0 = -(d / r) * kappa * M0 * M1 * c / rho[0]
    a1 = (d / r) * kappa * M1 * M0 * c / rho[1]
    got = _run2(_cfg2(kappa=kappa))
    np.testing.assert_allclose(got["accel"][0], a0, rtol=1e-5, atol=1e-6 * np.abs(a0).max())
    np.testing.assert_allclose(got["accel"][1], a1, rtol=1e-5, atol=1e-6 * np.abs(a1).max())
    assert float(got["accel"][0].astype(np.float64) @ d) < 0.0


def test_xsph_two_particle_closed_form_2d():
    nu = 0.3
    _, _, _, w1, rho, _ = _pair2()
    rb = 0.5 * (rho[0] + rho[1])
    v = VEL2.astype(np.float64)
    v0 = v[0] + nu * (v[1] - v[0]) * (M1 / rb) * w1
    v1 = v[1] + nu * (v[0] - v[1]) * (M0 / rb) * w1
    got = _run2(_cfg2(nu=nu), jacobi=True)
    np.testing.assert_allclose(got["vel"][0], v0, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(got["vel"][1], v1, rtol=1e-6, atol=1e-7)
    p0 = M0 * v[0] + M1 * v[1]
    p1 = M0 * got["vel"][0].astype(np.float64) + M1 * got["vel"][1].astype(np.float64)
    np.testing.assert_allclose(p1, p0, rtol=0, atol=1e-6)
    gs = _run2(_cfg2(nu=nu), jacobi=False)
    x0n = POS2[0].astype(np.float64) + v0 * 1e-4
    dn = x0n - POS2[1].astype(np.float64)
    v1_gs = v[1] + nu * (v0 - v[1]) * (M0 / rb) * _w2(float(dn @ dn))
    np.testing.assert_allclose(gs["vel"][1], v1_gs, rtol=1e-6, atol=1e-7)

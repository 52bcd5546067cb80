"""Synthetic dam-break scenes: seeded jittered lattices in static-box tanks.

Shared settings: smoothing length h = 0.1, lattice spacing s = 0.05, rest
density 1000, Tait gamma 7, particle mass chosen so the rest lattice sums to
the rest density (3D kernel sum over a spacing-h/2 lattice = 1.0097751668946855
/ s^3), XSPH 0.3, speed cap 30, particle radius s/2, gravity -9.81 along y,
walls 0.5 thick.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from .ffi import FluidConfig3Ffi

H = 0.1
SPACING = 0.05
RHO0 = 1000.0
LATTICE_SUM_3D = 1.0097751668946855  # sum_j W6(|x_j|) s^3 over the h = 2s lattice
MASS_3D = RHO0 * SPACING ** 3 / LATTICE_SUM_3D
WALL = 0.5


@dataclass
class Scene:
    name: str
    positions: np.ndarray  # (N, 3) float32
    config: FluidConfig3Ffi
    boxes: list = field(default_factory=list)  # [(center, half_extents)]
    gravity: tuple = (0.0, -9.81, 0.0)
    dt: float = 1e-4

    @property
    def n(self):
        return len(self.positions)


def fluid_config(stiffness, mass=MASS_3D, rho0=RHO0, h=H, gamma=7.0, viscosity=0.3,
                 radius=SPACING / 2):
    c = FluidConfig3Ffi.default()
    c.particle_mass = mass
    c.rest_density = rho0
    c.smoothing_length = h
    c.stiffness = stiffness
    c.gamma = gamma
    c.viscosity = viscosity
    c.max_velocity = 30.0
    c.particle_radius = radius
    for k in range(4):
        c.phase_rest_density[k] = rho0
        c.phase_mass[k] = mass
    return c


def tank_boxes(lx, ly, lz, wall_height=None):
    """Floor and four walls around the interior [0,lx] x [0,wall_height] x [0,lz]."""
    hy = wall_height if wall_height is not None else ly * 1.5
    t = WALL / 2
    return [
        ((lx / 2, -t, lz / 2), (lx / 2 + WALL, t, lz / 2 + WALL)),
        ((-t, hy / 2, lz / 2), (t, hy / 2, lz / 2 + WALL)),
        ((lx + t, hy / 2, lz / 2), (t, hy / 2, lz / 2 + WALL)),
        ((lx / 2, hy / 2, -t), (lx / 2 + WALL, hy / 2, t)),
        ((lx / 2, hy / 2, lz + t), (lx / 2 + WALL, hy / 2, t)),
    ]


def lattice(nx, ny, nz, origin=(0.0, 0.0, 0.0), spacing=SPACING, jitter=0.01, seed=0):
    """x-fastest lattice at cell centres, uniform +-jitter*spacing per axis."""
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pts = np.stack([ix.transpose(2, 1, 0).ravel(), iy.transpose(2, 1, 0).ravel(),
                    iz.transpose(2, 1, 0).ravel()], axis=1).astype(np.float64)
    pts = (pts + 0.5) * spacing + np.asarray(origin, dtype=np.float64)
    if jitter:
        rng = np.random.default_rng(seed)
        pts += rng.uniform(-jitter, jitter, size=pts.shape) * spacing
    return pts.astype(np.float32)


def stiffness_for_height(height, rho0=RHO0, gamma=7.0):
    """k with c_eff = sqrt(gamma k / rho0) = 10 sqrt(2 g H)."""
    c = 10.0 * math.sqrt(2 * 9.81 * height)
    return rho0 * c * c / gamma


def dam_break(nx, ny, nz, box=None, seed=2, jitter=0.01, dt=1e-4, name="dam"):
    """Column of nx x ny x nz particles against the x = 0 wall of a tank
    four times the column length along x."""
    lx, ly, lz = nx * SPACING, ny * SPACING, nz * SPACING
    bx, by, bz = box if box is not None else (4 * lx, ly, lz)
    pos = lattice(nx, ny, nz, seed=seed, jitter=jitter)
    k = stiffness_for_height(ly)
    return Scene(name, pos, fluid_config(k), tank_boxes(bx, by, bz), dt=dt)


def c2_dam_break_1m(seed=2):
    """1,048,576 particles: 64 x 128 x 128 column (3.2 x 6.4 x 6.4) in a
    12.8 x 6.4 x 6.4 tank (column = 1/4 of the volume), k = 1.8e6, dt = 1e-4."""
    s = dam_break(64, 128, 128, box=(12.8, 6.4, 6.4), seed=seed, name="dam_break_3d_1m")
    s.config.stiffness = 1.8e6
    return s


def c5_two_columns_8m(seed=5):
    """8,388,608 particles: two 64 x 256 x 256 columns at the x-ends of a
    12.8^3 tank, k = 3.6e6, dt = 1e-4."""
    a = lattice(64, 256, 256, seed=seed)
    b = lattice(64, 256, 256, origin=(12.8 - 64 * SPACING, 0.0, 0.0), seed=seed + 1000)
    pos = np.concatenate([a, b])
    return Scene("two_columns_3d_8m", pos, fluid_config(3.6e6), tank_boxes(12.8, 12.8, 12.8), dt=1e-4)


def small_dam(n_side=16, seed=7, jitter=0.01):
    """Small 3D dam (n_side x 2 n_side x n_side) with the production physics."""
    return dam_break(n_side, 2 * n_side, n_side, seed=seed, jitter=jitter, name=f"dam_{n_side}")


def batch_scenes(count=256, seed=3, nmin=2048, nmax=65536):
    """Independent dam scenes, particle counts log-uniform in [nmin, nmax],
    column aspect (height / length) uniform in [0.5, 2]."""
    rng = np.random.default_rng(seed)
    scenes = []
    for k in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        aspect = rng.uniform(0.5, 2.0)
        # nx * ny * nz ~ n with ny = aspect * nx and nz = nx
        nx = max(2, int(round((n / aspect) ** (1 / 3))))
        ny = max(2, int(round(aspect * nx)))
        nz = max(2, int(round(n / (nx * ny))))
        scenes.append(dam_break(nx, ny, nz, seed=seed * 100000 + k, name=f"batch_{k}"))
    return scenes


def make_world(scene, world_cls=None):
    """Create a World3 holding the scene (capacity lifted to its size)."""
    from .world import World3
    cls = world_cls or World3
    w = cls(capacity=max(scene.n, 4096))
    if not w.set_fluid_config(scene.config):
        raise ValueError("scene config rejected")
    w.set_gravity(scene.gravity)
    for c, he in scene.boxes:
        if w.add_static_box(c, he) < 0:
            raise ValueError("body rejected")
    got = w.add_particles(scene.positions)
    if got != scene.n:
        raise RuntimeError(f"added {got} of {scene.n} particles")
    return w


@dataclass
class Scene2:
    """A 2D scene: stepped on the GPU through the World2 exports
    (make_world2) and on the CPU by the 2D restatement."""
    name: str
    positions: np.ndarray  # (N, 2) float32
    config: dict           # FluidConfig2 fields
    boxes: list            # [(center (x, y), half_extents (hx, hy))]
    gravity: tuple = (0.0, -9.81)
    dt: float = 1.0 / 600.0
    steps: int = 1000

    @property
    def n(self):
        return len(self.positions)


def lattice_sum_2d(h=H, s=SPACING):
    """sum_j W2(|x_j|) over the spacing-s square lattice (the 2D poly6 kernel
    4 / (pi h^8) (h^2 - r^2)^3, fluid.rs:265-273), in f64."""
    k = int(math.ceil(h / s))
    return sum(4.0 / (math.pi * h ** 8) * max(h * h - (a * a + b * b) * s * s, 0.0) ** 3
               for a in range(-k, k + 1) for b in range(-k, k + 1))


def dam2d_config1(seed=1):
    """Config 1, the 2D dam break on the CPU: 4,096 particles on a 64 x 64
    lattice at s = h/2 = 0.05 (+-1% seeded jitter), a 3.2 x 3.2 column against
    the left wall of a 12.8-wide tank (walls 4.8 high), Tait gamma 7, rest
    density 1000, dt = 1/600, 1,000 steps.  The step divides the pressure
    force by rho a second time (fluid.rs:1006, acc = g + F / rho), so the
    stiffness carries a factor rho0: k = rho0 c^2 / gamma with c = 20 m/s
    (Courant number c dt / h = 0.33 without substeps)."""
    rng = np.random.default_rng(seed)
    ix, iy = np.meshgrid(np.arange(64), np.arange(64))
    pos = np.stack([ix.ravel(), iy.ravel()], axis=1).astype(np.float64)
    pos = (pos + 0.5) * SPACING + rng.uniform(-0.01, 0.01, size=pos.shape) * SPACING
    mass = RHO0 / lattice_sum_2d()
    cfg = dict(rest_density=RHO0, smoothing_length=H, stiffness=RHO0 * 20.0 ** 2 / 7.0, gamma=7.0,
               viscosity=0.3, max_velocity=30.0, particle_radius=SPACING / 2, mass=mass)
    t = WALL / 2
    boxes = [((6.4, -t), (6.4 + WALL, t)), ((-t, 2.4), (t, 2.4)), ((12.8 + t, 2.4), (t, 2.4))]
    return Scene2("dam_break_2d_4096", pos.astype(np.float32), cfg, boxes)


def scene2_config(sc):
    """FluidConfig2Ffi (FluidConfig3Ffi's layout) of a Scene2: the defaults
    with the scene's fields, one phase of mass c["mass"]."""
    from .ffi import FluidConfig3Ffi
    c = sc.config
    cfg = FluidConfig3Ffi.default()
    cfg.particle_mass = c["mass"]
    cfg.rest_density = c["rest_density"]
    cfg.smoothing_length = c["smoothing_length"]
    cfg.stiffness = c["stiffness"]
    cfg.gamma = c["gamma"]
    cfg.viscosity = c["viscosity"]
    cfg.max_velocity = c["max_velocity"]
    cfg.particle_radius = c["particle_radius"]
    for k in range(4):
        cfg.phase_rest_density[k] = c["rest_density"]
        cfg.phase_mass[k] = c["mass"]
    return cfg


def make_world2(sc, capacity=None):
    """A World2 holding the Scene2 (static boxes in the scene's order)."""
    from .world import World2
    w = World2(capacity=capacity or max(sc.n, 4096))
    if not w.set_fluid_config(scene2_config(sc)):
        raise ValueError("scene config rejected")
    w.set_gravity(sc.gravity)
    for c, he in sc.boxes:
        if w.add_static_box(c, he) < 0:
            raise ValueError("body rejected")
    got = w.add_particles(sc.positions)
    if got != sc.n:
        raise RuntimeError(f"added {got} of {sc.n} particles")
    return w


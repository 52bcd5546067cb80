"""Write the golden fixtures of one Jacobi substep from the CPU oracle.

The reference ships no numeric vectors for its fluid path, so these are the
oracle's outputs on small seeded scenes (inputs and outputs both stored).
The oracle itself is pinned by tests/test_oracle.py; these fixtures freeze
its arithmetic and give the GPU tests inputs that do not depend on the
scene generator.

    python -m tests.golden.make_golden
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from phi_eng_amd import scenes  # noqa: E402
from phi_eng_amd.ffi import FluidConfig3Ffi  # noqa: E402


def _bodies_array(boxes=(), spheres=(), capsules=()):
    rows = [(1, *c, *he) for c, he in boxes]
    rows += [(0, *c, r, 0.0, 0.0) for c, r in spheres]
    rows += [(2, *c, r, hl, 0.0) for c, r, hl in capsules]
    return np.array(rows, dtype=np.float64).reshape(-1, 7)


def _oracle_bodies(arr):
    boxes = [((r[1], r[2], r[3]), (r[4], r[5], r[6])) for r in arr if int(r[0]) == 1]
    spheres = [((r[1], r[2], r[3]), r[4]) for r in arr if int(r[0]) == 0]
    caps = [((r[1], r[2], r[3]), r[4], r[5]) for r in arr if int(r[0]) == 2]
    # keep the stored order: boxes/spheres/capsules are written in that order
    return O.bodies_from_boxes(boxes, spheres, caps)


def config_from(g):
    return FluidConfig3Ffi.from_buffer_copy(bytes(g["config"]))


def recompute(g):
    cfg = config_from(g)
    r = O.substep_f32(g["pos_in"], g["vel_in"], cfg, g["gravity"], _oracle_bodies(g["bodies"]),
                      float(g["dt"]), jacobi=True, phases=g["phase"])
    return dict(density=r["density"], pressure=r["pressure"], accel=r["accel"], count=r["count"],
                pos_out=r["pos"], vel_out=r["vel"], scale=r["scale"])


def cases():
    out = {}
    sc = scenes.small_dam(8, seed=101)
    out["dam8_rest"] = dict(pos=sc.positions, vel=np.zeros_like(sc.positions), cfg=sc.config,
                            g=sc.gravity, bodies=_bodies_array(sc.boxes), dt=sc.dt, phase=None)
    sc = scenes.small_dam(6, seed=102)
    rng = np.random.default_rng(102)
    cfg = sc.config.copy()
    cfg.artificial_viscosity = 0.8
    cfg.surface_tension = 0.5
    cfg.wall_friction = 0.3
    pos = sc.positions + np.float32(0.02)
    bodies = _bodies_array(sc.boxes, spheres=[((0.15, 0.2, 0.15), 0.06)], capsules=[((0.2, 0.3, 0.1), 0.04, 0.1)])
    ph = rng.integers(0, 6, size=len(pos)).astype(np.uint8)
    cfg.phase_rest_density[1] = 900.0
    cfg.phase_rest_density[2] = 1100.0
    cfg.phase_mass[1] = cfg.phase_mass[0] * 0.9
    out["dam6_options"] = dict(pos=pos, vel=rng.normal(0, 0.5, size=pos.shape).astype(np.float32), cfg=cfg,
                               g=(0.0, -9.81, 0.0), bodies=bodies, dt=5e-4, phase=ph)
    # edge cases: coincident pair (r < EPS), an isolated particle, one deep
    # inside a wall, particles at negative cells, default (linear) config
    cfg = FluidConfig3Ffi.default()
    pos = np.array([[0.1, 0.1, 0.1], [0.1, 0.1, 0.1], [0.3, 0.1, 0.1], [5.0, 5.0, 5.0],
                    [-0.3, -0.2, -0.1], [-0.1, -0.2, -0.1], [1.2, -0.95, 0.5]], dtype=np.float32)
    out["edge_default_cfg"] = dict(pos=pos, vel=np.zeros_like(pos), cfg=cfg, g=(0.0, -9.81, 0.0),
                                   bodies=_bodies_array([((1.0, -1.0, 0.5), (3.0, 0.5, 3.0))]), dt=1 / 60,
                                   phase=None)
    return out


def main():
    for name, c in cases().items():
        n = len(c["pos"])
        phase = c["phase"] if c["phase"] is not None else np.zeros(n, dtype=np.uint8)
        g = dict(pos_in=c["pos"].astype(np.float32), vel_in=c["vel"].astype(np.float32),
                 config=np.frombuffer(bytes(c["cfg"]), dtype=np.uint8), gravity=np.asarray(c["g"], np.float32),
                 bodies=c["bodies"], dt=np.float32(c["dt"]), phase=phase)
        out = recompute(g)
        g.update({k: v for k, v in out.items() if k != "scale"})
        g["scale"] = out["scale"]
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **g)
        print(name, n)


if __name__ == "__main__":
    main()

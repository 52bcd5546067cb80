"""GPU parity of one fluid substep against the f32 CPU oracle.

Contract (crates/dynamics/src/fluid3.rs restated in oracle/sph_oracle.cpp):
cell key triples, sorted order and neighbour counts bit-exact; density,
pressure and acceleration within 1e-5 relative (acceleration against the
cancellation-aware scale max(|a|, |g| + sum_j |m_j term_ij grad W_ij|/rho_i));
the Jacobi integrate and boundary clamp within 1e-5 of the oracle's Jacobi
mode.  The kernels keep the reference operation order, so most of these
come out bit-identical; the tests state the tolerance and report the max.
"""
import numpy as np
import pytest

from phi_eng_amd import scenes
from phi_eng_amd.world import RECORD_ACCEL, RECORD_CELLS, RECORD_DENSITY

pytestmark = pytest.mark.gpu

REL = 1e-5


def _rel(a, b, floor=1e-30):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def _compare_step(w, scene, O, label):
    pos0, vel0 = w.get_state()
    bodies = O.bodies_from_boxes(scene.boxes)
    ref = O.substep_f32(pos0, vel0, scene.config, scene.gravity, bodies, scene.dt, jacobi=True)
    w.set_record(RECORD_CELLS | RECORD_DENSITY | RECORD_ACCEL)
    w.step(scene.dt)
    rec = w.get_record()
    pos1, vel1 = w.get_state()
    h = scene.config.smoothing_length
    cells = O.cells(pos0, h)
    order = O.sorted_ids(pos0, h)
    print(f"[{label}] n={len(pos0)}", flush=True)
    assert np.array_equal(rec["cells"], cells), "cell keys differ"
    assert np.array_equal(rec["order"], order), "sorted order differs"
    assert np.array_equal(rec["count"], ref["count"]), "neighbour counts differ"
    rd = _rel(rec["density"], ref["density"]).max()
    rp = np.abs(rec["pressure"].astype(np.float64) - ref["pressure"]).max() / max(1.0, np.abs(ref["pressure"]).max())
    g = np.linalg.norm(scene.gravity)
    s = np.maximum(np.linalg.norm(ref["accel"], axis=1), g + ref["scale"])
    ra = (np.abs(rec["accel"].astype(np.float64) - ref["accel"]).max(axis=1) / s).max()
    dx = np.abs(pos1.astype(np.float64) - ref["pos"]).max()
    vs = np.maximum(np.abs(ref["vel"]).max(), 1.0)
    dv = np.abs(vel1.astype(np.float64) - ref["vel"]).max() / vs
    bit_rho = np.mean(rec["density"].view(np.uint32) == ref["density"].view(np.uint32))
    bit_acc = np.mean(rec["accel"].view(np.uint32) == ref["accel"].view(np.uint32))
    print(f"[{label}] rho rel {rd:.3e} (bit-equal {bit_rho:.4f})  p {rp:.3e}  acc {ra:.3e} "
          f"(bit-equal {bit_acc:.4f})  pos abs {dx:.3e}  vel rel {dv:.3e}", flush=True)
    assert rd <= REL
    assert rp <= REL
    assert ra <= REL
    assert dx <= REL * max(1.0, np.abs(ref["pos"]).max())
    assert dv <= REL
    return ref


def test_first_step_small_dam(gpu, oracle_lib):
    sc = scenes.small_dam(16, seed=11)
    w = scenes.make_world(sc)
    _compare_step(w, sc, oracle_lib, "dam16 step1")


def test_evolved_state_with_velocity(gpu, oracle_lib):
    sc = scenes.small_dam(12, seed=12)
    sc.dt = 5e-4
    w = scenes.make_world(sc)
    for _ in range(30):
        w.step(sc.dt)
    _compare_step(w, sc, oracle_lib, "dam12 evolved")


def test_random_velocities_av_cohesion(gpu, oracle_lib):
    sc = scenes.small_dam(10, seed=13)
    sc.config.artificial_viscosity = 0.5
    sc.config.surface_tension = 0.2
    w = scenes.make_world(sc)
    rng = np.random.default_rng(5)
    w.set_velocity(rng.normal(0, 0.5, size=(sc.n, 3)).astype(np.float32))
    _compare_step(w, sc, oracle_lib, "dam10 av+coh")


def test_outliers_and_negative_cells(gpu, oracle_lib):
    """Particles far outside the grid domain and at negative cell indices."""
    sc = scenes.small_dam(8, seed=14)
    far = np.array([[40.0, 3.0, 0.2], [40.05, 3.02, 0.21], [-7.3, -2.0, -5.5], [-7.35, -2.01, -5.52]],
                   dtype=np.float32)
    sc.positions = np.concatenate([sc.positions - np.float32(0.4), far])
    w = scenes.make_world(sc)
    _compare_step(w, sc, oracle_lib, "outliers")


def test_cfl_substeps_follow_oracle(gpu, oracle_lib):
    """CFL substepping (fluid3.rs:591-636) on the GPU: the substep count comes
    from the device speed maximum, each substep is dt / n; three steps of
    the world against the oracle's Jacobi run of the same steps."""
    O = oracle_lib
    sc = scenes.small_dam(10, seed=16)
    sc.config.stiffness = 0.02  # c0 = sqrt(7 k rho0) ~ 11.8 m/s: a few substeps at dt = 6e-3
    sc.config.cfl = 0.3
    sc.config.max_substeps = 8
    dt = 6e-3
    w = scenes.make_world(sc)
    pos0, vel0 = w.get_state()
    for _ in range(3):
        w.step(dt)
    pos, vel = w.get_state()
    rp, rv, sub = O.run_f32(pos0, vel0, sc.config, sc.gravity, O.bodies_from_boxes(sc.boxes), dt, 3, jacobi=True)
    dx = np.abs(pos.astype(np.float64) - rp).max()
    dv = np.abs(vel.astype(np.float64) - rv).max() / max(np.abs(rv).max(), 1.0)
    print(f"cfl: oracle substeps in the last step {sub}, max |dx| {dx:.3e}, vel rel {dv:.3e}, "
          f"bit-equal {np.mean(pos.view(np.uint32) == rp.view(np.uint32)):.4f}", flush=True)
    assert 2 <= sub <= 8
    assert dx <= REL * max(1.0, np.abs(rp).max()) and dv <= REL
    w.close()


def test_add_particles_after_steps(gpu, oracle_lib):
    """Particles added to a world that has stepped (the tile count grows
    from 7 to 8 and to 9 on the second add), then the fluid cleared and a
    new set added (3 tiles): every step equals the oracle's Jacobi substep
    of the same particle set bit for bit."""
    from phi_eng_amd.ffi import FluidConfig3Ffi
    from phi_eng_amd.world import World3
    O = oracle_lib
    cfg = FluidConfig3Ffi.default()
    boxes = [((2.0, -1.0, 2.0), (3.0, 0.5, 3.0)), ((-1.0, 2.0, 2.0), (0.5, 3.5, 3.5)),
             ((5.0, 2.0, 2.0), (0.5, 3.5, 3.5)), ((2.0, 2.0, -1.0), (3.5, 3.5, 0.5)),
             ((2.0, 2.0, 5.0), (3.5, 3.5, 0.5))]
    bodies = O.bodies_from_boxes(boxes)
    rng = np.random.default_rng(21)

    def block(n_side, lo):
        g = np.arange(n_side) * 0.3
        p = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + np.asarray(lo)
        return (p + rng.uniform(-0.02, 0.02, p.shape)).astype(np.float32)

    w = World3()
    assert w.set_fluid_config(cfg)
    for c, he in boxes:
        assert w.add_static_box(c, he) >= 0
    rp = block(12, (0.2, 0.2, 0.2))
    rv = np.zeros_like(rp)
    assert w.add_particles(rp) == len(rp)
    dt = 1.0 / 60.0
    plan = ((block(6, (2.3, 1.0, 2.3)), 0), (None, 5), (block(7, (0.4, 2.5, 0.4)), 0), (None, 5), ("clear", 0),
            (block(9, (1.0, 0.3, 1.0)), 4))
    for more, steps in plan:
        if isinstance(more, str):
            w.clear()
            rp = np.zeros((0, 3), np.float32)
            rv = np.zeros((0, 3), np.float32)
        elif more is not None:
            assert w.add_particles(more) == len(more)
            rp = np.concatenate([rp, more])
            rv = np.concatenate([rv, np.zeros_like(more)])
        for k in range(steps):
            w.step(dt)
            r = O.substep_f32(rp, rv, cfg, (0.0, -9.81, 0.0), bodies, dt, jacobi=True)
            rp, rv = r["pos"], r["vel"]
            p, v = w.get_state()
            print(f"[add] {len(p)} particles, step {k}", flush=True)
            assert np.array_equal(p.view(np.uint32), rp.view(np.uint32)), f"positions, {len(p)} particles"
            assert np.array_equal(v.view(np.uint32), rv.view(np.uint32)), f"velocities, {len(p)} particles"
    w.close()


def test_run_twice_bit_identical(gpu):
    """fluid3.rs:1398-1416 determinism, on the GPU path."""
    sc = scenes.small_dam(10, seed=15)
    outs = []
    for _ in range(2):
        w = scenes.make_world(sc)
        for _ in range(20):
            w.step(5e-4)
        outs.append(w.get_state())
        w.close()
    assert np.array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
    assert np.array_equal(outs[0][1].view(np.uint32), outs[1][1].view(np.uint32))


def _world_from_fixture(g):
    from phi_eng_amd.world import World3
    from tests.golden.make_golden import config_from
    w = World3(capacity=max(4096, len(g["pos_in"])))
    assert w.set_fluid_config(config_from(g))
    w.set_gravity(g["gravity"])
    for row in g["bodies"]:
        assert w.add_body(row[1:4], int(row[0]), row[4], row[5], row[6]) >= 0
    assert w.add_particles(g["pos_in"], phases=g["phase"]) == len(g["pos_in"])
    assert w.set_velocity(g["vel_in"]) == len(g["pos_in"])
    return w


@pytest.mark.parametrize("name", ["dam8_rest", "dam6_options", "edge_default_cfg"])
def test_golden_fixture(gpu, name):
    """One GPU substep from each committed fixture equals the fixture's
    oracle outputs (counts exactly; floats within 1e-5, bit-equality reported)."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", f"{name}.npz"))
    w = _world_from_fixture(g)
    w.set_record(RECORD_CELLS | RECORD_DENSITY | RECORD_ACCEL)
    w.step(float(g["dt"]))
    rec = w.get_record()
    pos1, vel1 = w.get_state()
    print(f"[{name}] n={len(pos1)}", flush=True)
    assert np.array_equal(rec["count"], g["count"])
    assert _rel(rec["density"], g["density"]).max() <= REL
    s = np.maximum(np.linalg.norm(g["accel"], axis=1), np.linalg.norm(g["gravity"]) + g["scale"])
    assert (np.abs(rec["accel"].astype(np.float64) - g["accel"]).max(axis=1) / s).max() <= REL
    assert np.abs(pos1.astype(np.float64) - g["pos_out"]).max() <= REL * max(1.0, np.abs(g["pos_out"]).max())
    assert np.abs(vel1.astype(np.float64) - g["vel_out"]).max() <= REL * max(1.0, np.abs(g["vel_out"]).max())
    bits = np.mean(np.concatenate([rec["density"].view(np.uint32) == g["density"].view(np.uint32),
                                   (rec["accel"].view(np.uint32) == g["accel"].view(np.uint32)).ravel(),
                                   (pos1.view(np.uint32) == g["pos_out"].view(np.uint32)).ravel()]))
    print(f"[{name}] bit-equal fraction {bits:.4f}", flush=True)


def test_full_size_1m_step(gpu, oracle_lib):
    """The 1,048,576-particle dam break: one step against the oracle at full
    size (cells, order, counts exact; floats within 1e-5), after 20 GPU steps."""
    sc = scenes.c2_dam_break_1m()
    w = scenes.make_world(sc)
    for _ in range(20):
        w.step(sc.dt)
    _compare_step(w, sc, oracle_lib, "1M after 20 steps")


def test_full_size_invariants(gpu):
    """Size-independent properties at 1M over 100 steps: count preserved,
    finite, contained by the tank, mass conserved, bounded speed."""
    sc = scenes.c2_dam_break_1m()
    w = scenes.make_world(sc)
    for _ in range(100):
        w.step(sc.dt)
    pos, vel = w.get_state()
    assert len(pos) == sc.n == w.count()
    assert np.isfinite(pos).all() and np.isfinite(vel).all()
    pr = sc.config.particle_radius
    assert (pos[:, 0] >= pr - 1e-4).all() and (pos[:, 0] <= 12.8 - pr + 1e-4).all()
    assert (pos[:, 1] >= pr - 1e-4).all()
    assert (pos[:, 2] >= pr - 1e-4).all() and (pos[:, 2] <= 6.4 - pr + 1e-4).all()
    assert np.linalg.norm(vel, axis=1).max() <= sc.config.max_velocity * (1 + 1e-6)

"""Adaptive resolution (config adaptive_resolution; fluid3.rs:923-987 and
1007-1012, fluid.rs:826-891) on the GPU against the oracle: the particle
set after each substep (count, order, positions, velocities, masses) is
bit-identical to the oracle's Jacobi step, merges and splits included; the
reference's adaptive tests run on the GPU path."""
import numpy as np
import pytest

from oracle import oracle as O
from phi_eng_amd.ffi import FluidConfig3Ffi
from phi_eng_amd.world import RECORD_ACCEL, RECORD_DENSITY, Batch3, Surrogate3, World2, World3

pytestmark = pytest.mark.gpu

F = np.float32
TANK3 = [((2.0, -1.0, 2.0), (3.0, 0.5, 3.0)), ((-1.0, 2.0, 2.0), (0.5, 3.5, 3.5)),
         ((5.0, 2.0, 2.0), (0.5, 3.5, 3.5)), ((2.0, 2.0, -1.0), (3.5, 3.5, 0.5)),
         ((2.0, 2.0, 5.0), (3.5, 3.5, 0.5))]
TANK2 = [((2.0, -1.0), (3.0, 0.5)), ((-1.0, 2.0), (0.5, 3.5)), ((5.0, 2.0), (0.5, 3.5))]


@pytest.fixture(autouse=True)
def _oracle_built(oracle_lib):
    return oracle_lib


def _bits_equal(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                                                 np.ascontiguousarray(b).view(np.uint32))


def _mixed_scene(seed, n_block=6):
    """A dense block (spacing 0.2, density ~4 rest: every particle above the
    merge ratio) with pairs pushed inside h/4, plus isolated thin particles
    (density ~0.4 rest: below the split ratio)."""
    rng = np.random.default_rng(seed)
    g = np.arange(n_block) * 0.2 + 1.0
    blk = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    blk += rng.uniform(-0.01, 0.01, blk.shape)
    close = rng.choice(len(blk), len(blk) // 6, replace=False)
    pair = blk[close] + rng.normal(0, 0.05, (len(close), 3))
    iso = np.stack([rng.uniform(0.0, 4.0, 40), rng.uniform(2.5, 3.0, 40), rng.uniform(0.0, 4.0, 40)], -1)
    iso = iso[np.argsort(iso[:, 0])]
    keep = [0]
    for k in range(1, len(iso)):
        if np.min(np.linalg.norm(iso[keep] - iso[k], axis=1)) > 0.6:
            keep.append(k)
    pos = np.concatenate([blk, pair, iso[keep]]).astype(F)
    return pos[rng.permutation(len(pos))]


def _cfg(**kw):
    c = FluidConfig3Ffi.default()
    c.adaptive_resolution = 1
    c.phase_mass[0] = 0.03
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _world3(cfg, pos, vel=None, boxes=TANK3):
    w = World3()
    assert w.set_fluid_config(cfg)
    for c, he in boxes:
        assert w.add_static_box(c, he) >= 0
    assert w.add_particles(pos) == len(pos)
    if vel is not None:
        assert w.set_velocity(vel) == len(pos)
    return w


def _state(w):
    p, v = w.get_state()
    m, ph = w.get_mass()
    return p, v, m, ph


def _compare(w, ref, label):
    p, v, m, _ = _state(w)
    print(f"[{label}] gpu {len(p)} particles, oracle {len(ref['pos'])}", flush=True)
    assert len(p) == len(ref["pos"]), f"{label}: particle count"
    assert _bits_equal(m, ref["mass"]), f"{label}: masses"
    assert _bits_equal(p, ref["pos"]), f"{label}: positions"
    assert _bits_equal(v, ref["vel"]), f"{label}: velocities"


def test_one_substep_with_merges_and_splits():
    """One substep of a scene where both passes act: the particle set, the
    densities of the new set and the acceleration match the oracle bit for
    bit."""
    cfg = _cfg()
    pos = _mixed_scene(3)
    rng = np.random.default_rng(4)
    vel = rng.normal(0, 0.5, pos.shape).astype(F)
    w = _world3(cfg, pos, vel)
    mass = O.masses(cfg, len(pos))
    ref = O.substep_cap_f32(pos, vel, mass, cfg, (0.0, -9.81, 0.0), O.bodies_from_boxes(TANK3), 1.0 / 60.0, 4096)
    assert len(ref["pos"]) != len(pos)
    assert (ref["mass"] > F(0.03)).any() and (ref["mass"] < F(0.03)).any()  # merges and splits happened
    w.set_record(RECORD_DENSITY | RECORD_ACCEL)
    w.step(1.0 / 60.0)
    w.sync()
    rec = w.get_record()
    _compare(w, ref, "substep 1")
    assert np.array_equal(rec["count"], ref["count"])
    assert _bits_equal(rec["density"], ref["density"])
    assert _bits_equal(rec["accel"], ref["accel"])
    w.close()


def test_trajectory_3d_tank():
    """40 steps of the mixed scene in the reference tank: the set evolves
    (merges, splits, removal) exactly as the oracle's."""
    cfg = _cfg()
    pos = _mixed_scene(5)
    w = _world3(cfg, pos)
    for _ in range(40):
        w.step(1.0 / 60.0)
    w.sync()
    ref = O.run_cap_f32(pos, np.zeros_like(pos), O.masses(cfg, len(pos)), cfg, (0.0, -9.81, 0.0),
                        O.bodies_from_boxes(TANK3), 1.0 / 60.0, 40, 4096)
    _compare(w, ref, "40 steps")
    w.close()


def test_reference_adaptive_conserves_mass_2d():
    """fluid.rs:1519-1549 on the GPU: calm water in the 2D tank, 200 frames;
    mass within 1e-3, count within the cap, contained; and the state equals
    the oracle's Jacobi run."""
    w = World2()
    c = w.get_fluid_config()
    c.adaptive_resolution = 1
    assert w.set_fluid_config(c)
    for ctr, he in TANK2:
        assert w.add_static_box(ctr, he) >= 0
    pos = np.array([(-0.5 + ix * 0.25, -0.5 + iy * 0.25) for iy in range(6) for ix in range(12)], F)
    assert w.add_particles(pos) == len(pos)
    m0 = float(w.get_mass()[0].sum())
    for _ in range(200):
        w.step(1.0 / 60.0)
    w.sync()
    p, v = w.get_state()
    m, _ = w.get_mass()
    assert np.isfinite(p).all() and np.isfinite(v).all()
    assert abs(float(m.astype(np.float64).sum()) - m0) < 1e-3
    assert len(p) <= 4096
    assert (p[:, 0] > -1.0).all() and (p[:, 0] < 5.0).all() and (p[:, 1] > -0.6).all()
    cfg2 = O.Oracle2Cfg.default()
    cfg2.adaptive_resolution = 1
    ref = O.run2_cap_f32(pos, None, np.full(len(pos), 0.0625, F), cfg2, (0.0, -9.81), O.bodies2(boxes=TANK2),
                         1.0 / 60.0, 200, 4096)
    print(f"[2D adaptive] gpu {len(p)}, oracle {len(ref['pos'])}", flush=True)
    assert len(p) == len(ref["pos"])
    assert _bits_equal(m, ref["mass"])
    assert _bits_equal(p, ref["pos"]) and _bits_equal(v, ref["vel"])
    w.close()


def test_2d_merges_and_splits_follow_oracle():
    """The 2D pass (fluid.rs:826-891) with both merges and splits: a dense
    block (spacing 0.15, ~2.8 rest) with pairs inside h/4 and isolated thin
    particles, 30 steps in the 2D tank."""
    rng = np.random.default_rng(12)
    g = np.arange(10) * 0.15 + 0.5
    blk = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    close = rng.choice(len(blk), 20, replace=False)
    iso = np.stack([np.arange(6) * 0.7 + 0.2, np.full(6, 3.0)], -1)
    pos = np.concatenate([blk, blk[close] + rng.normal(0, 0.04, (20, 2)), iso]).astype(F)
    pos = pos[rng.permutation(len(pos))]
    w = World2()
    c = w.get_fluid_config()
    c.adaptive_resolution = 1
    assert w.set_fluid_config(c)
    for ctr, he in TANK2:
        assert w.add_static_box(ctr, he) >= 0
    assert w.add_particles(pos) == len(pos)
    for _ in range(30):
        w.step(1.0 / 60.0)
    w.sync()
    p, v = w.get_state()
    m, _ = w.get_mass()
    cfg2 = O.Oracle2Cfg.default()
    cfg2.adaptive_resolution = 1
    ref = O.run2_cap_f32(pos, None, np.full(len(pos), 0.0625, F), cfg2, (0.0, -9.81), O.bodies2(boxes=TANK2),
                         1.0 / 60.0, 30, 4096)
    print(f"[2D mixed] gpu {len(p)}, oracle {len(ref['pos'])} (from {len(pos)})", flush=True)
    assert (ref["mass"] > F(0.0625)).any() and (ref["mass"] < F(0.0625)).any()
    assert len(p) == len(ref["pos"])
    assert _bits_equal(m, ref["mass"])
    assert _bits_equal(p, ref["pos"]) and _bits_equal(v, ref["vel"])
    w.close()


def test_split_cap_at_capacity():
    """Near the capacity (4,096) the split cap of fluid3.rs:972 decides which
    thin particles split; the count may pass the cap as the reference's
    does."""
    rng = np.random.default_rng(9)
    g = np.arange(16) * 0.6
    pos = np.stack(np.meshgrid(g, g, g[:15], indexing="ij"), -1).reshape(-1, 3)[:4000]
    pos = (pos + rng.uniform(-0.05, 0.05, pos.shape)).astype(F)
    cfg = _cfg()
    w = _world3(cfg, pos, boxes=[])
    w.step(1.0 / 60.0)
    w.sync()
    ref = O.substep_cap_f32(pos, np.zeros_like(pos), O.masses(cfg, len(pos)), cfg, (0.0, -9.81, 0.0),
                            O.bodies_from_boxes([]), 1.0 / 60.0, 4096)
    assert 4096 < len(ref["pos"]) < 2 * len(pos)
    _compare(w, ref, "cap")
    w.close()


def test_adaptive_refuses_surrogate_and_batch():
    cfg = _cfg()
    w = _world3(cfg, _mixed_scene(1, 3))
    net = Surrogate3(seed=2)
    with pytest.raises(RuntimeError):
        w.set_surrogate(net)
    with pytest.raises(RuntimeError):
        Batch3([w])
    net.close()
    w.close()
    w = World3()
    w.add_particles(np.zeros((1, 3), F))
    net = Surrogate3(seed=2)
    w.set_surrogate(net)
    assert not w.set_fluid_config(_cfg())
    net.close()
    w.close()


def test_adaptive_with_ghosts_and_cfl_substeps():
    """Adaptive resolution together with boundary ghosts (fluid3.rs:1014-1027
    runs after the resample) and CFL substeps (a changing count between
    substeps of one step): 10 steps equal the oracle's."""
    cfg = _cfg(boundary_particles=1, cfl=0.3, max_substeps=4, stiffness=400.0)
    cfg.particle_mass = cfg.phase_mass[0]
    pos = _mixed_scene(8)
    w = _world3(cfg, pos)
    for _ in range(10):
        w.step(1.0 / 60.0)
    w.sync()
    assert w.boundary_info()[0] > 0
    ref = O.run_cap_f32(pos, np.zeros_like(pos), O.masses(cfg, len(pos)), cfg, (0.0, -9.81, 0.0),
                        O.bodies_from_boxes(TANK3), 1.0 / 60.0, 10, 4096)
    assert (ref["mass"] != F(0.03)).any()
    _compare(w, ref, "adaptive + ghosts + cfl")
    w.close()


def test_candidate_store_regrow_keeps_parity(monkeypatch):
    """The merge-candidate store starts at 4 entries (test hook), so the
    first merge pass overflows it: that pass must write nothing, the host
    grows the store and runs the pass again, and the substep still equals
    the oracle bit for bit (the particle arrays are untouched by the
    overflowed pass)."""
    monkeypatch.setenv("PHIENG_RESAMPLE_STORE_INITIAL", "4")
    cfg = _cfg()
    pos = _mixed_scene(3)
    rng = np.random.default_rng(4)
    vel = rng.normal(0, 0.5, pos.shape).astype(F)
    w = _world3(cfg, pos, vel)
    mass = O.masses(cfg, len(pos))
    ref = O.substep_cap_f32(pos, vel, mass, cfg, (0.0, -9.81, 0.0), O.bodies_from_boxes(TANK3), 1.0 / 60.0, 4096)
    assert (ref["mass"] > F(0.03)).sum() > 4  # more merges than the first store holds
    w.step(1.0 / 60.0)
    w.sync()
    _compare(w, ref, "regrown store")
    w.close()

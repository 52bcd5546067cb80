"""Verlet neighbour lists on the plain substep: walk substeps equal the
oracle's Jacobi substep bit for bit (fluid3.rs:639-655 density,
:1029-1098 forces, in the reference's neighbour order), whether the rows
are walked, put back in order, or rebuilt; and the list can be switched
off or given any skin without changing a single bit."""
import numpy as np
import pytest

from phi_eng_amd import scenes
from phi_eng_amd.world import RECORD_ACCEL, RECORD_DENSITY

pytestmark = pytest.mark.gpu


def _stirred_dam(n_side, seed, speed):
    """A small dam launched at `speed` m/s along (2, 0.5, 1) / |.| plus
    random velocities of up to speed / 4, so that many particles cross cell
    faces between rebuilds (random motion alone is smoothed away by the
    viscosity within a few substeps)."""
    sc = scenes.small_dam(n_side, seed=seed)
    w = scenes.make_world(sc)
    w.set_neighbour_list(True, 0.1)  # the row form at any size (the default takes it from 131,072 up)
    rng = np.random.default_rng(seed)
    pos, _ = w.get_state()
    d = np.array([2.0, 0.5, 1.0]) / np.linalg.norm([2.0, 0.5, 1.0])
    vel = (speed * d + rng.uniform(-speed / 4, speed / 4, pos.shape)).astype(np.float32)
    w.set_velocity(vel)
    return sc, w


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_walk_trajectory_bit_exact(gpu, oracle_lib):
    """40 substeps of a stirred dam, each against the oracle's Jacobi
    substep of the same state: positions and velocities bit-equal; most
    substeps walk the rows and some rows are put back in order."""
    O = oracle_lib
    sc, w = _stirred_dam(10, seed=31, speed=3.0)
    dt = 2e-4
    bodies = O.bodies_from_boxes(sc.boxes)
    rp, rv = w.get_state()
    for k in range(40):
        w.step(dt)
        r = O.substep_f32(rp, rv, sc.config, sc.gravity, bodies, dt, jacobi=True)
        rp, rv = r["pos"], r["vel"]
        p, v = w.get_state()
        print(f"[verlet walk] step {k}", flush=True)
        assert np.array_equal(_bits(p), _bits(rp)), f"positions differ at step {k}"
        assert np.array_equal(_bits(v), _bits(rv)), f"velocities differ at step {k}"
    st = w.neighbour_stats()
    print(f"[verlet walk] stats {st}", flush=True)
    assert st["substeps"] + st["list_substeps"] == 40
    assert 1 <= st["rebuilds"] < st["substeps"]
    assert st["repaired"] > 0
    w.close()


def test_walk_records_bit_exact(gpu, oracle_lib):
    """Density, pressure, neighbour count and acceleration records of walk
    substeps (no cell record, which would sort) equal the oracle's."""
    O = oracle_lib
    sc, w = _stirred_dam(8, seed=32, speed=1.0)
    dt = 2e-4
    bodies = O.bodies_from_boxes(sc.boxes)
    for _ in range(3):
        w.step(dt)
    w.set_record(RECORD_DENSITY | RECORD_ACCEL)
    for k in range(5):
        p0, v0 = w.get_state()
        ref = O.substep_f32(p0, v0, sc.config, sc.gravity, bodies, dt, jacobi=True)
        w.step(dt)
        rec = w.get_record()
        print(f"[verlet records] step {k}", flush=True)
        assert np.array_equal(rec["count"], ref["count"])
        assert np.array_equal(_bits(rec["density"]), _bits(ref["density"]))
        assert np.array_equal(_bits(rec["pressure"]), _bits(ref["pressure"]))
        assert np.array_equal(_bits(rec["accel"]), _bits(ref["accel"]))
    st = w.neighbour_stats()
    assert st["rebuilds"] < st["substeps"]
    w.close()


@pytest.mark.parametrize("verlet,skin", [(False, 0.1), (True, 0.0), (True, 0.3), ("auto", 0.1)])
def test_settings_do_not_change_results(gpu, verlet, skin):
    """The list off, a zero skin (a rebuild every substep), a wide skin and
    the automatic choice (the cells, at this size) give the row-form
    world's trajectory bit for bit."""
    dt = 2e-4
    _, a = _stirred_dam(10, seed=33, speed=1.5)
    _, b = _stirred_dam(10, seed=33, speed=1.5)
    b.set_neighbour_list(verlet, skin)
    for k in range(30):
        a.step(dt)
        b.step(dt)
    pa, va = a.get_state()
    pb, vb = b.get_state()
    sa, sb = a.neighbour_stats(), b.neighbour_stats()
    print(f"[verlet settings] verlet={verlet} skin={skin} default {sa} other {sb}", flush=True)
    assert np.array_equal(_bits(pa), _bits(pb))
    assert np.array_equal(_bits(va), _bits(vb))
    if verlet and skin == 0.0:
        assert sb["rebuilds"] == sb["substeps"] == 30
    a.close()
    b.close()


def test_fast_flow_switches_to_the_list_form(gpu):
    """A dam stirred fast enough that more than 1,024 rows in a substep need
    re-ordering: the world sweeps the cells for a stretch of substeps, then
    rebuilds the rows and walks again, and its trajectory is the list-off
    world's bit for bit throughout."""
    dt = 2e-4
    _, a = _stirred_dam(16, seed=34, speed=6.0)
    _, b = _stirred_dam(16, seed=34, speed=6.0)
    b.set_neighbour_list(False, 0.1)
    for k in range(60):
        a.step(dt)
        b.step(dt)
        if k % 20 == 19:
            pa, va = a.get_state()
            pb, vb = b.get_state()
            print(f"[verlet switch] step {k}: {a.neighbour_stats()}", flush=True)
            assert np.array_equal(_bits(pa), _bits(pb)), f"positions differ by step {k}"
            assert np.array_equal(_bits(va), _bits(vb)), f"velocities differ by step {k}"
    st = a.neighbour_stats()
    assert st["list_substeps"] > 0 and st["substeps"] > 0
    assert st["substeps"] + st["list_substeps"] == 60
    a.close()
    b.close()

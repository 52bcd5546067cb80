"""Batched stepping (no reference counterpart): every scene of a batch must
be bit-identical to the same world stepped alone through the plain path,
in every launch mode, and autotuning must not change results."""
import numpy as np
import pytest

from phi_eng_amd import scenes
from phi_eng_amd.world import BATCH_KERNELS, BATCH_KERNELS_GRAPH, BATCH_PERSISTENT, BATCH_PERSISTENT_GRAPH, Batch3

pytestmark = pytest.mark.gpu

STEPS = 12
DT = 2e-4


def _scenes(k=6, seed=41):
    return scenes.batch_scenes(count=k, seed=seed, nmin=300, nmax=6000)


def _alone(scs):
    out = []
    for sc in scs:
        w = scenes.make_world(sc)
        for _ in range(STEPS):
            w.step(DT)
        out.append(w.get_state())
        w.close()
    return out


def _batched(scs, mode, bpc=0):
    ws = [scenes.make_world(sc) for sc in scs]
    b = Batch3(ws)
    b.set_mode(mode, bpc)
    for _ in range(STEPS):
        b.step(DT)
    b.sync()
    b.sync_worlds()
    out = [w.get_state() for w in ws]
    b.close()
    return out


def _same(a, b):
    return all(np.array_equal(x[0].view(np.uint32), y[0].view(np.uint32)) and
               np.array_equal(x[1].view(np.uint32), y[1].view(np.uint32)) for x, y in zip(a, b))


def test_batch_matches_worlds_alone_all_modes(gpu):
    scs = _scenes()
    ref = _alone(scs)
    for mode, bpc in [(BATCH_KERNELS, 0), (BATCH_PERSISTENT, 0), (BATCH_PERSISTENT, 1),
                      (BATCH_PERSISTENT_GRAPH, 0), (BATCH_PERSISTENT_GRAPH, 2), (BATCH_KERNELS_GRAPH, 0)]:
        got = _batched(scs, mode, bpc)
        print(f"mode {mode} blocks/CU {bpc}: n={sum(len(g[0]) for g in got)}", flush=True)
        assert _same(got, ref), (mode, bpc)


def test_autotune_keeps_state_and_results(gpu, tmp_path):
    scs = _scenes(5, seed=43)
    ref = _alone(scs)
    ws = [scenes.make_world(sc) for sc in scs]
    b = Batch3(ws)
    cache = str(tmp_path / "tune.txt")
    mode, bpc, ms = b.autotune(DT, trial_steps=4, cache_path=cache)
    print(f"autotune picked mode {mode} blocks/CU {bpc} ({ms:.3f} ms/step)", flush=True)
    assert mode in (BATCH_KERNELS, BATCH_KERNELS_GRAPH, BATCH_PERSISTENT_GRAPH)
    for _ in range(STEPS):
        b.step(DT)
    b.sync()
    b.sync_worlds()
    assert _same([w.get_state() for w in ws], ref)
    # the cache answers the same shape without trials
    m2, bpc2, _ = b.autotune(DT, trial_steps=4, cache_path=cache)
    assert (m2, bpc2) == (mode, bpc)
    b.close()


def test_batch_rejects_cfl_worlds(gpu):
    sc = _scenes(1)[0]
    sc.config.cfl = 0.4
    sc.config.max_substeps = 4
    w = scenes.make_world(sc)
    with pytest.raises(RuntimeError, match="CFL"):
        Batch3([w])


def test_graph_replays_match_eager_every_step(gpu):
    """The multi-kernel substep captured in a hipGraph, replayed step after
    step with the host's own work (readback, monitor) between replays:
    every replay's state equals the eager multi-kernel batch's."""
    scs = scenes.batch_scenes(count=6, seed=9, nmin=1000, nmax=6000)
    eager_w = [scenes.make_world(sc) for sc in scs]
    graph_w = [scenes.make_world(sc) for sc in scs]
    eager, graph = Batch3(eager_w), Batch3(graph_w)
    eager.set_mode(BATCH_KERNELS, 0)
    graph.set_mode(BATCH_KERNELS_GRAPH, 0)
    for step in range(20):
        eager.step(DT)
        graph.step(DT)
        eager.sync()
        graph.sync()
        eager.sync_worlds()
        graph.sync_worlds()
        a = [w.get_state() for w in eager_w]
        b = [w.get_state() for w in graph_w]
        assert _same(a, b), f"replay {step} differs"
    eager.close()
    graph.close()


def test_barrier_timeout_is_reported_and_recovered(gpu):
    """The persistent launch's grid barrier waits a bounded number of spins;
    with the limit at zero a block that has to wait gives up at once, every
    block leaves at its next barrier and the launch ends normally.  The next
    sync reports it, steps are refused until set_mode clears the barrier,
    and the batch then steps again."""
    scs = _scenes(4, seed=47)
    ws = [scenes.make_world(sc) for sc in scs]
    b = Batch3(ws)
    b.set_mode(BATCH_PERSISTENT, 0)
    b.set_spin_limit(0)
    for _ in range(3):
        b.step(DT)
    with pytest.raises(RuntimeError, match="timed out"):
        b.sync()
    with pytest.raises(RuntimeError, match="timed out earlier"):
        b.step(DT)
    b.set_spin_limit(1 << 22)
    b.set_mode(BATCH_PERSISTENT, 0)
    for _ in range(3):
        b.step(DT)
    b.sync()
    b.close()
    for w in ws:
        w.close()


def test_mode_switches_after_verlet_substeps(gpu):
    """A batch above the automatic Verlet threshold (131,072 particles), so
    its eager substeps walk Verlet rows and alternate the state generation:
    eager for an odd number of steps, then hipGraph replays (captured on the
    generation the eager steps left), then eager again.  After every phase
    each scene equals the world stepped alone, bit for bit."""
    scs = scenes.batch_scenes(count=6, seed=12, nmin=24000, nmax=32000)
    assert sum(sc.n for sc in scs) >= 131072
    alone = [scenes.make_world(sc) for sc in scs]
    ws = [scenes.make_world(sc) for sc in scs]
    b = Batch3(ws)
    for phase, (mode, steps) in enumerate([(BATCH_KERNELS, 7), (BATCH_KERNELS_GRAPH, 6), (BATCH_KERNELS, 5)]):
        b.set_mode(mode, 0)
        for _ in range(steps):
            b.step(DT)
            for w in alone:
                w.step(DT)
        b.sync()
        b.sync_worlds()
        got = [w.get_state() for w in ws]
        ref = [w.get_state() for w in alone]
        print(f"[batch switch] phase {phase} mode {mode}: {sum(len(g[0]) for g in got)} particles", flush=True)
        assert _same(got, ref), f"phase {phase} (mode {mode}) differs"
    b.close()
    for w in alone:
        w.close()

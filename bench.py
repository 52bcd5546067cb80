#!/usr/bin/env python3
"""Benchmark: plain 3D SPH dam break, 1,048,576 particles, on one MI355X.

Prints one JSON line (particle-steps/s over the timed steps, device-resident
state, synthetic seeded input), with the HBM roofline of the dominant kernel
from HIP-event kernel timings, the FP64 error of one GPU substep, and the
CPU oracle timed on a bounded sample of the same scene: on all host cores
(OpenMP Jacobi, cpu_baseline) and on one core (the reference's Gauss-Seidel
form, cpu_baseline_1core).

    python bench.py [--gpus 1] [--steps 500] [--warmup 20] [--sur-steps 200]

The "surrogate" object holds the learned-pressure form on the same scene:
its throughput, speed-up over the plain path over the same steps, and the
mean density error against the plain run, with the network-free ablation.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/sec at 1M-particle 3D SPH; surrogate vs plain at ≤1% density err"
HBM_PEAK_GBS = 8000.0
# Compulsory bytes per particle of each pass (f32 SoA; every array read and
# written once; neighbour re-reads not counted).
ALG_BYTES = {"cell_key": 20, "scan": 0, "scatter": 16, "rank_gather": 68, "density": 20,
             "force_integrate": 56}
B_ALG_STEP = 185
# VALU issue peak of the MI355X: 256 CUs x 4 SIMDs, one wave instruction per
# 4 cycles each, at the 2.4 GHz peak engine clock (wave instructions / s).
VALU_PEAK_WIPS = 1024 * 2.4e9 / 4


def _args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=2,
                    help="sample length (steps) of the one-core CPU baseline (the reference's Gauss-Seidel form)")
    ap.add_argument("--cpu-par-steps", type=int, default=10,
                    help="sample length (steps) of the all-cores CPU baseline (OpenMP Jacobi)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="HBM bytes per launch per kernel from rocprofv3 FETCH_SIZE/WRITE_SIZE (optional)")
    ap.add_argument("--valu", default=os.path.join(ROOT, "profiles", "valu.json"),
                    help="VALU wave instructions per launch per kernel from rocprofv3 SQ_INSTS_VALU (optional)")
    ap.add_argument("--workload", default="plain", choices=["plain", "batch", "scale8m", "extras", "dam2d"],
                    help="plain: the 1M dam break (the headline); batch: 256 uneven scenes in one launch per step; "
                         "scale8m: 8,388,608 particles, two colliding columns, plain vs surrogate; extras: the "
                         "1M dam break with adaptive resolution, and snapshot save / load of the 1M world; "
                         "dam2d: the 4,096-particle 2D dam break alone (GPU and CPU)")
    ap.add_argument("--no-2d", action="store_true",
                    help="skip the 2D dam break timed beside the plain workload (profiled runs)")
    ap.add_argument("--calibrate", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (rocprof counter scale)")
    ap.add_argument("--sur-steps", type=int, default=200,
                    help="steps of the surrogate vs plain comparison on the 1M scene (0 skips it)")
    ap.add_argument("--sur-train-steps", type=int, default=2000,
                    help="steps of the plain training trajectory the network is trained on")
    return ap.parse_args()


# Plain judge every CHECK_EVERY substeps in the surrogate runs (the bench
# re-measures 15 and 20 beside it).  Each judge also evaluates the surrogate
# on 1/16 of the particles of the judged state and falls back above an
# acceleration error sum |a_plain - a_sur| / sum |a_plain| of 0.2, besides
# the bulk density deviation of 1%.
CHECK_EVERY = 10


def bench_surrogate(a, plain_value):
    """The learned pressure term on the 1M scene: train the network on GPU
    from a plain GPU trajectory of a half-scale dam (samples every 10 steps
    from 1/16 of the particles), then run the 1M scene from the same start
    with plain SPH, with the surrogate (plain judge every CHECK_EVERY steps,
    fallback above 1% bulk density deviation), and with the surrogate's network
    switched off; density error = mean |rho - rho_plain| / rho_plain after
    the run."""
    from phi_eng_amd import scenes

    net, train = train_network(a.sur_train_steps)
    sc = scenes.c2_dam_break_1m()
    n = sc.n
    S = a.sur_steps
    sweep = ("surrogate_ce15", "surrogate_ce20")
    runs = compare_runs(sc, net, S, ("plain", "surrogate", "network_free") + sweep)
    ref = runs["plain"]["rho"]
    out = {}
    for name in ("surrogate", "network_free") + sweep:
        r = runs[name]
        e = np.abs(r["rho"] - ref) / np.maximum(ref, 1e-6)
        out[name] = {"value": round(n * S / r["wall"], 1), "ms_per_step": round(r["wall"] * 1e3 / S, 4),
                     "speedup_vs_plain": round(runs["plain"]["wall"] / r["wall"], 4),
                     "density_err_mean": float(e.mean()), "density_err_p99": float(np.percentile(e, 99)),
                     "surrogate_steps": r["stats"]["surrogate_steps"], "plain_steps": r["stats"]["plain_steps"],
                     "fallbacks": r["stats"]["fallbacks"], "judge_accel_err": r["judge"][1],
                     "judge_density_err": r["judge"][0], "kept_judges": r["judge"][2]}
    net.close()
    res = out["surrogate"]
    res.update({
        "plain_ms_per_step_same_steps": round(runs["plain"]["wall"] * 1e3 / S, 4),
        "speedup_vs_plain_headline": round(res["value"] / plain_value, 4),
        "steps": S, "check_every": CHECK_EVERY, "max_density_error": 0.01, "max_accel_error": 0.2,
        "judge": "every check_every-th substep plain: bulk |rho/rho0 - 1| and, over 1/16 of the particles, "
                 "sum |a_plain - a_sur| / sum |a_plain| of the surrogate evaluated on the judged state",
        "skin": "Verlet: max(0.01, 4 k check_every dt vmax / h), k = 1 or 2 judge intervals per list",
        "network_free": out["network_free"],
        "check_every_sweep": {k.split("_ce")[1]: {"speedup_vs_plain": out[k]["speedup_vs_plain"],
                                                   "density_err_mean": out[k]["density_err_mean"],
                                                   "fallbacks": out[k]["fallbacks"]} for k in sweep},
        "train": train,
        "data": "synthetic: seeded jittered lattice dam breaks; network trained in this run",
    })
    return res


def fp64_error(w, sc):
    """Max and p99 over all particles of one GPU substep's density (relative)
    and acceleration (relative to max(|a|, |g| + sum_j |m_j term grad W_ij| /
    rho_i), the cancellation-aware scale) against the FP64 twin of the same
    substep from the same state."""
    from oracle import oracle as O
    from phi_eng_amd.world import RECORD_ACCEL, RECORD_DENSITY
    pos, vel = w.get_state()
    w.set_record(RECORD_DENSITY | RECORD_ACCEL)
    w.step(sc.dt)
    rec = w.get_record()
    w.set_record(0)
    t0 = time.perf_counter()
    r = O.substep_f64(pos, vel, sc.config, sc.gravity, O.bodies_from_boxes(sc.boxes), sc.dt)
    t1 = time.perf_counter()
    drho = np.abs(rec["density"].astype(np.float64) - r["density"]) / np.maximum(np.abs(r["density"]), 1e-30)
    g = float(np.linalg.norm(sc.gravity))
    scale = np.maximum(np.linalg.norm(r["accel"], axis=1), g + r["scale"])
    dacc = np.abs(rec["accel"].astype(np.float64) - r["accel"]).max(axis=1) / scale
    return {"density_rel_max": float(drho.max()), "density_rel_p99": float(np.percentile(drho, 99)),
            "accel_rel_max": float(dacc.max()), "accel_rel_p99": float(np.percentile(dacc, 99)),
            "particles": int(len(pos)), "twin_seconds": round(t1 - t0, 2),
            "what": "one substep of the 1M dam break from the state after the timed run; acceleration error "
                    "relative to max(|a|, |g| + sum |m term grad W| / rho)"}


def bench_cpu_2d():
    """The 2D dam break (4,096 particles, 1,000 steps at dt 1/600): the
    reference's 2D step is CPU code, timed here as its restatement on one
    host thread."""
    from oracle import oracle as O
    from phi_eng_amd import scenes

    sc = scenes.dam2d_config1()
    cfg, mass, bodies = O.scene2_inputs(sc)
    t0 = time.perf_counter()
    out = O.run2_f32(sc.positions, np.zeros_like(sc.positions), mass, cfg, sc.gravity, bodies, sc.dt, sc.steps)
    wall = time.perf_counter() - t0
    d = out["density"]
    return {"workload": sc.name, "particles": sc.n, "steps": sc.steps, "dt": sc.dt,
            "value": round(sc.n * sc.steps / wall, 1), "unit": "particle-steps/s", "wall_s": round(wall, 3),
            "cores": 1, "kind": "port",
            "sample": "the whole workload: f32 restatement of fluid.rs (in-place integrate), 1 thread",
            "final_density_median": float(np.median(d)),
            "contained": bool(np.isfinite(out["pos"]).all() and (out["pos"][:, 0] > 0).all()
                              and (out["pos"][:, 0] < 12.8).all() and (out["pos"][:, 1] > 0).all())}


def bench_gpu_2d():
    """The same 2D dam break on the GPU through the World2 exports (the 3D
    engine's kernels on the z = 0 plane with the 2D constants; Jacobi
    integrate), whole workload: 1,000 steps of 4,096 particles."""
    from phi_eng_amd import scenes

    sc = scenes.dam2d_config1()
    w = scenes.make_world2(sc)
    for _ in range(5):
        w.step(sc.dt)
    w.sync()
    w.close()
    w = scenes.make_world2(sc)
    w.sync()
    t0 = time.perf_counter()
    for _ in range(sc.steps):
        w.step(sc.dt)
    w.sync()
    wall = time.perf_counter() - t0
    pos, _ = w.get_state()
    w.close()
    return {"workload": sc.name, "particles": sc.n, "steps": sc.steps, "dt": sc.dt,
            "value": round(sc.n * sc.steps / wall, 1), "unit": "particle-steps/s", "wall_s": round(wall, 4),
            "ms_per_step": round(wall * 1e3 / sc.steps, 4),
            "contained": bool(np.isfinite(pos).all() and (pos[:, 0] > 0).all() and (pos[:, 0] < 12.8).all()
                              and (pos[:, 1] > 0).all())}


def train_network(steps):
    """Collect samples from a plain GPU trajectory of a half-scale dam (every
    10 steps, 1/16 of the particles) and train the network on GPU."""
    from phi_eng_amd import scenes
    from phi_eng_amd.world import Surrogate3

    t0 = time.perf_counter()
    tr = scenes.dam_break(32, 64, 64, box=(6.4, 3.2, 3.2), seed=4, name="train")
    tr.config.stiffness = 1.8e6
    wt = scenes.make_world(tr)
    net = Surrogate3(seed=4)
    wt.collect_samples(net, every=10, stride=16)
    for _ in range(steps):
        wt.step(tr.dt)
    wt.sync()
    wt.close()
    t1 = time.perf_counter()
    loss0, _ = net.evaluate()
    first, last, base = net.train(epochs=8, batch=4096, lr=2e-3, seed=1)
    t2 = time.perf_counter()
    return net, {"scene": "dam 32x64x64 (131072 particles), 6.4x3.2x3.2 tank, k 1.8e6",
                 "trajectory_steps": steps, "samples": int(net.num_samples()),
                 "epochs": 8, "loss_untrained": round(loss0, 6), "loss_network_free": round(base, 6),
                 "loss_first_epoch": round(first, 6), "loss_last_epoch": round(last, 6),
                 "collect_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2)}


def compare_runs(sc, net, steps, names):
    """Run the scene from its start with each form (plain, surrogate,
    network_free) for `steps` timed steps after one untimed set-up step;
    return wall time, surrogate statistics and the density after the run."""
    from phi_eng_amd import scenes
    from phi_eng_amd.world import RECORD_DENSITY

    runs = {}
    for name in names:
        w = scenes.make_world(sc)
        if name != "plain":
            ce = int(name.split("_ce")[1]) if "_ce" in name else CHECK_EVERY
            w.set_surrogate(net, check_every=ce, max_density_error=0.01, use_network=not name.startswith("network_free"))
        w.step(sc.dt)  # first step (set-up, judge) outside the timing
        w.sync()
        ta = time.perf_counter()
        for _ in range(steps):
            w.step(sc.dt)
        w.sync()
        tb = time.perf_counter()
        st = w.surrogate_stats()
        judge = w.surrogate_judge() if name != "plain" else None
        w.set_surrogate(None)
        w.set_record(RECORD_DENSITY)
        w.step(sc.dt)
        rho = w.get_record()["density"].astype(np.float64)
        runs[name] = dict(wall=tb - ta, stats=st, rho=rho, judge=judge)
        if name == "plain":
            runs[name]["state"] = w.get_state()
        w.close()
        print(f"{name}: {steps} steps in {tb - ta:.3f} s", file=sys.stderr, flush=True)
    return runs


def bench_scale(a):
    """8,388,608 particles (two 64 x 256 x 256 columns at the x-ends of a
    12.8^3 tank, k 3.6e6, dt 1e-4): plain GPU, surrogate in the loop (plain
    judge every CHECK_EVERY steps), and the CPU oracle on a bounded sample."""
    from phi_eng_amd import scenes

    steps = a.steps if a.steps != 500 else 300
    net, train = train_network(a.sur_train_steps)
    sc = scenes.c5_two_columns_8m()
    n = sc.n
    runs = compare_runs(sc, net, steps, ("plain", "surrogate"))
    ref = runs["plain"]["rho"]
    e = np.abs(runs["surrogate"]["rho"] - ref) / np.maximum(ref, 1e-6)
    pv = n * steps / runs["plain"]["wall"]
    sv = n * steps / runs["surrogate"]["wall"]
    cpu = None
    if not a.no_cpu and a.cpu_steps > 0:
        from oracle import oracle as O
        pos, vel = runs["plain"]["state"]
        bodies = O.bodies_from_boxes(sc.boxes)
        c0 = time.perf_counter()
        O.run_f32(pos, vel, sc.config, sc.gravity, bodies, sc.dt, 1, jacobi=False)
        c1 = time.perf_counter()
        cpu = {"value": round(n / (c1 - c0), 1), "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": f"1 step of the same {n}-particle scene (state after the plain GPU run), "
                         "f32 hash-grid Gauss-Seidel restatement of fluid3.rs, 1 thread"}
    st = runs["surrogate"]["stats"]
    out = {"metric": "particle-steps/sec at 8.4M-particle 3D SPH; surrogate vs plain", "value": round(sv, 1),
           "unit": "particle-steps/s", "n_gpus": 1, "steps": steps, "warmup": 1,
           "ms_per_step": round(runs["surrogate"]["wall"] * 1e3 / steps, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic: two seeded jittered lattice columns; network trained in this run",
           "config": {"workload": "two_columns_3d_8m", "particles": n, "lattice": [[64, 256, 256], [64, 256, 256]],
                      "tank": [12.8, 12.8, 12.8], "dt": sc.dt, "stiffness": sc.config.stiffness,
                      "path": "surrogate", "check_every": CHECK_EVERY, "max_density_error": 0.01,
                      "max_accel_error": 0.2, "skin": "Verlet, min 0.01"},
           "plain": {"value": round(pv, 1), "ms_per_step": round(runs["plain"]["wall"] * 1e3 / steps, 4)},
           "speedup_vs_plain": round(sv / pv, 4), "density_err_mean": float(e.mean()),
           "density_err_p99": float(np.percentile(e, 99)),
           "surrogate_steps": st["surrogate_steps"], "plain_steps": st["plain_steps"], "fallbacks": st["fallbacks"],
           "judge_accel_err": runs["surrogate"]["judge"][1], "kept_judges": runs["surrogate"]["judge"][2],
           "train": train, "cpu_baseline": cpu}
    print(json.dumps(out))


def calibrate(mib):
    """A plain copy of known size through the HIP runtime (its copy kernel
    reads and writes mib MiB), for scaling FETCH_SIZE / WRITE_SIZE."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(5):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    print(json.dumps({"calibrate_copy_bytes": n, "copies": 5}))


def main():
    a = _args()
    if a.calibrate:
        calibrate(a.calibrate)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU form of this "
              "path is design-only", file=sys.stderr)
        sys.exit(2)
    if a.workload == "batch":
        return bench_batch(a)
    if a.workload == "scale8m":
        return bench_scale(a)
    if a.workload == "extras":
        return bench_extras(a)
    if a.workload == "dam2d":
        return bench_dam2d(a)
    from phi_eng_amd import scenes
    from phi_eng_amd.ffi import last_error

    sc = scenes.c2_dam_break_1m()
    w = scenes.make_world(sc)
    n = sc.n
    for _ in range(a.warmup):
        w.step(sc.dt)
    w.sync()

    t0 = time.perf_counter()
    for _ in range(a.steps):
        w.step(sc.dt)
    w.sync()
    t1 = time.perf_counter()
    err = last_error()
    if err:
        raise RuntimeError(err)
    wall = t1 - t0
    value = n * a.steps / wall
    ms_per_step = wall * 1e3 / a.steps

    nstats = w.neighbour_stats()
    # Per-kernel HIP-event timings over a second run of the same steps from
    # the same start (the flow, and so the neighbour search's form, as in the
    # timed run).
    wt = scenes.make_world(sc)
    for _ in range(a.warmup):
        wt.step(sc.dt)
    wt.sync()
    wt.set_timing(True)
    t2 = time.perf_counter()
    for _ in range(a.steps):
        wt.step(sc.dt)
    wt.sync()
    t3 = time.perf_counter()
    timing = wt.get_timing()
    wt.close()
    kernels = {k: {"avg_us": 1e3 * ms / cnt, "launches": int(cnt)} for k, (ms, cnt) in timing.items()}
    dom = max((k for k in kernels if k in ALG_BYTES and ALG_BYTES[k]), key=lambda k: kernels[k]["avg_us"])
    avg_s = kernels[dom]["avg_us"] * 1e-6
    achieved = ALG_BYTES[dom] * n / avg_s / 1e9
    traffic = None
    if os.path.exists(a.traffic):
        with open(a.traffic) as f:
            tr = json.load(f)
        traffic = tr.get("bytes_per_launch", {}).get(dom)
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "alg_bytes_per_particle": ALG_BYTES[dom], "avg_launch_us": round(kernels[dom]["avg_us"], 2)}
    # The stencil passes are VALU-issue-bound, not HBM-bound: the same kernel
    # against the VALU issue peak, its instruction count from rocprofv3.
    if os.path.exists(a.valu):
        with open(a.valu) as f:
            vi = json.load(f).get("per_launch", {}).get(dom)
        if vi:
            wips = vi["insts_valu"] / avg_s
            roof["valu"] = {"insts_per_launch": vi["insts_valu"], "achieved": round(wips / 1e9, 1),
                            "peak": round(VALU_PEAK_WIPS / 1e9, 1), "unit": "G wave-instructions/s",
                            "frac": round(wips / VALU_PEAK_WIPS, 4)}
    step_frac = B_ALG_STEP * n / (wall / a.steps) / (HBM_PEAK_GBS * 1e9)

    # FP64 error of the GPU result: one recorded substep from the current
    # state against the oracle's FP64 twin of the same substep.
    fp64 = None if a.no_cpu else fp64_error(w, sc)

    cpu = cpu1 = None
    if not a.no_cpu:
        from oracle import oracle as O
        bodies = O.bodies_from_boxes(sc.boxes)
        pos, vel = w.get_state()
        if a.cpu_par_steps > 0:
            threads = O.max_threads()
            c0 = time.perf_counter()
            O.run_f32_par(pos, vel, sc.config, sc.gravity, bodies, sc.dt, a.cpu_par_steps, threads=threads)
            c1 = time.perf_counter()
            cpu = {"value": round(n * a.cpu_par_steps / (c1 - c0), 1), "unit": "particle-steps/s",
                   "cores": threads, "nproc": os.cpu_count(), "kind": "port",
                   "sample": f"{a.cpu_par_steps} steps of the same {n}-particle scene (state after the GPU run), "
                             f"the oracle's f32 Jacobi restatement of fluid3.rs with its per-particle loops on "
                             f"{threads} OpenMP threads (grid rebuild on one)"}
        if a.cpu_steps > 0:
            c0 = time.perf_counter()
            O.run_f32(pos, vel, sc.config, sc.gravity, bodies, sc.dt, a.cpu_steps, jacobi=False)
            c1 = time.perf_counter()
            cpu1 = {"value": round(n * a.cpu_steps / (c1 - c0), 1), "unit": "particle-steps/s", "cores": 1,
                    "kind": "port",
                    "sample": f"{a.cpu_steps} steps of the same {n}-particle scene (state after the GPU run), "
                              "f32 hash-grid Gauss-Seidel restatement of fluid3.rs, 1 thread"}

    cpu2d = None
    if not a.no_cpu and not a.no_2d:
        cpu2d = bench_cpu_2d()
    gpu2d = None if a.no_2d else bench_gpu_2d()

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded jittered lattice dam break",
        "config": {"workload": "dam_break_3d_1m", "particles": n, "lattice": [64, 128, 128],
                   "tank": [12.8, 6.4, 6.4], "h": 0.1, "spacing": 0.05, "dt": sc.dt,
                   "stiffness": sc.config.stiffness, "gamma": 7.0, "path": "plain"},
        "roofline": roof,
        "step_roofline_frac": round(step_frac, 5),
        "kernels": {k: round(v["avg_us"], 2) for k, v in kernels.items()},
        "timed_with_events_ms_per_step": round((t3 - t2) * 1e3 / a.steps, 4),
        "neighbour_list": dict(nstats, kind="verlet", skin=0.1),
        "cpu_baseline": cpu,
        "cpu_baseline_1core": cpu1,
        "fp64_err": fp64,
        "cpu_dam_break_2d": cpu2d,
        "gpu_dam_break_2d": gpu2d,
    }
    w.close()
    if a.sur_steps > 0:
        out["surrogate"] = bench_surrogate(a, value)
    print(json.dumps(out))


def bench_dam2d(a):
    """The 2D dam break (4,096 particles, 1,000 steps at dt 1/600) alone: the
    whole workload on the GPU, and on one CPU core unless --no-cpu."""
    gpu = bench_gpu_2d()
    line = {"metric": "particle-steps/sec, 2D dam break (4,096 particles, 1,000 steps)", "value": gpu["value"],
            "unit": "particle-steps/s", "n_gpus": 1, "steps": gpu["steps"], "warmup": 5,
            "ms_per_step": gpu["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic: seeded jittered 64 x 64 lattice",
            "config": {"workload": gpu["workload"], "particles": gpu["particles"], "dt": gpu["dt"]},
            "contained": gpu["contained"]}
    if not a.no_cpu:
        line["cpu_baseline"] = bench_cpu_2d()
    print(json.dumps(line), flush=True)


def bench_extras(a):
    """The 1M dam break with adaptive resolution (capacity 4M, so splits are
    not capped; split ratio 0.6, so the thinnest surface particles, whose
    density starts at ~0.51 rest, split from the first substep) against the
    same run without it; then JSON and binary snapshots of the 1M world:
    write, read back, and check the loaded state."""
    import tempfile

    import numpy as np

    from phi_eng_amd import scenes
    from phi_eng_amd.ffi import last_error
    from phi_eng_amd.world import World3

    sc = scenes.c2_dam_break_1m()
    steps = max(1, min(a.steps, 200))
    runs = {}
    for name, adaptive in (("plain", False), ("adaptive", True)):
        w = scenes.make_world(sc)
        w.reserve(4 * sc.n)
        cfg = w.get_fluid_config()
        cfg.adaptive_resolution = int(adaptive)
        cfg.split_density_ratio = 0.6
        cfg.particle_mass = cfg.phase_mass[0]
        assert w.set_fluid_config(cfg)
        for _ in range(5):
            w.step(sc.dt)
        w.sync()
        counts = 0
        t0 = time.perf_counter()
        for k in range(steps):
            counts += w.count()
            w.step(sc.dt)
        w.sync()
        wall = time.perf_counter() - t0
        err = last_error()
        if err:
            raise RuntimeError(err)
        m, _ = w.get_mass()
        runs[name] = {"value": round(counts / wall, 1), "ms_per_step": round(wall * 1e3 / steps, 4),
                      "particles_end": int(w.count()), "mass_total": float(m.astype(np.float64).sum())}
        if name == "adaptive":
            snap = {}
            with tempfile.TemporaryDirectory() as d:
                for fmt, binary in (("json", False), ("binary", True)):
                    path = os.path.join(d, "w." + fmt)
                    t0 = time.perf_counter()
                    rc = w.save(path, binary=binary)
                    t1 = time.perf_counter()
                    if rc != 1:
                        raise RuntimeError(f"save {fmt}: {rc} {last_error()}")
                    r = World3.load(path)
                    t2 = time.perf_counter()
                    if r is None:
                        raise RuntimeError(f"load {fmt}: {last_error()}")
                    p0, v0 = w.get_state()
                    p1, v1 = r.get_state()
                    same = bool(np.array_equal(p0.view(np.uint32), p1.view(np.uint32))
                                and np.array_equal(v0.view(np.uint32), v1.view(np.uint32)))
                    r.close()
                    mb = os.path.getsize(path) / 1e6
                    snap[fmt] = {"bytes": os.path.getsize(path), "save_s": round(t1 - t0, 3),
                                 "load_s": round(t2 - t1, 3), "save_MB_s": round(mb / (t1 - t0), 1),
                                 "load_MB_s": round(mb / (t2 - t1), 1), "state_identical": same}
            runs[name]["snapshot"] = snap
        w.close()
    ad, pl = runs["adaptive"], runs["plain"]
    out = {"metric": "particle-steps/sec at 1M-particle 3D SPH with adaptive resolution", "value": ad["value"],
           "unit": "particle-steps/s", "n_gpus": 1, "steps": steps, "warmup": 5, "ms_per_step": ad["ms_per_step"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic: seeded jittered lattice dam break",
           "config": {"workload": "dam_break_3d_1m_adaptive", "particles": sc.n, "capacity": 4 * sc.n,
                      "split_density_ratio": 0.6, "merge_density_ratio": 1.6, "dt": sc.dt,
                      "path": "plain + resample"},
           "plain": pl, "adaptive": ad,
           "cost_vs_plain": round(ad["ms_per_step"] / pl["ms_per_step"], 3)}
    print(json.dumps(out))


BATCH_MODES = {0: "kernels", 1: "persistent", 2: "persistent+hipGraph", 3: "kernels+hipGraph"}


def bench_batch(a):
    """256 independent dam scenes (counts log-uniform in [2048, 65536], seed 3)
    advanced together; the launch form (eight launches or one persistent
    launch per substep, each with or without a hipGraph) and the persistent
    grid are autotuned, and every form is timed."""
    from phi_eng_amd import scenes
    from phi_eng_amd.world import Batch3
    scs = scenes.batch_scenes(count=256, seed=3)
    worlds = [scenes.make_world(sc) for sc in scs]
    b = Batch3(worlds)
    n = b.count()
    dt = 1e-4
    cache = os.environ.get("PHIENG_TUNE_CACHE", os.path.join(tempfile.gettempdir(), "phieng_batch_tune_cache.txt"))
    mode, bpc, tuned_ms = b.autotune(dt, trial_steps=10, cache_path=cache)
    for _ in range(a.warmup):
        b.step(dt)
    b.sync()
    steps = a.steps if a.steps != 500 else 200
    t0 = time.perf_counter()
    for _ in range(steps):
        b.step(dt)
    b.sync()
    wall = time.perf_counter() - t0
    b.set_timing(True)
    for _ in range(20):
        b.step(dt)
    b.sync()
    timing = b.get_timing()
    b.set_timing(False)
    kern = {k: round(1e3 * ms / c, 2) for k, (ms, c) in timing.items()}
    # every launch form, each from the state it finds (50 timed steps; the
    # persistent forms at the occupancy limit of blocks per CU)
    forms = {}
    for m, name in BATCH_MODES.items():
        b.set_mode(m, 0)
        for _ in range(3):
            b.step(dt)
        b.sync()
        f0 = time.perf_counter()
        for _ in range(50):
            b.step(dt)
        b.sync()
        f_ms = (time.perf_counter() - f0) * 1e3 / 50
        forms[name] = {"ms_per_step": round(f_ms, 4), "value": round(n / (f_ms * 1e-3), 1)}
    b.set_mode(mode, bpc)
    out = {"metric": "particle-steps/sec, batched 3D SPH (256 scenes)", "value": round(n * steps / wall, 1),
           "unit": "particle-steps/s", "n_gpus": 1, "steps": steps, "warmup": a.warmup,
           "ms_per_step": round(wall * 1e3 / steps, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic: 256 seeded jittered dam breaks",
           "config": {"workload": "batch_256_scenes", "particles": n, "scenes": len(scs),
                      "counts": [int(min(sc.n for sc in scs)), int(max(sc.n for sc in scs))], "dt": dt,
                      "mode": BATCH_MODES[mode],
                      "blocks_per_cu": bpc, "autotune_ms_per_step": round(tuned_ms, 4)},
           "step_roofline_frac": round(B_ALG_STEP * n / (wall / steps) / (HBM_PEAK_GBS * 1e9), 5),
           "kernels": kern, "launch_forms": forms}
    print(json.dumps(out))


if __name__ == "__main__":
    main()

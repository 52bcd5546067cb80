"""Summarise rocprofv3 counter runs of the 1M dam break into the files
bench.py reads (profiles/traffic.json, profiles/valu.json).

Each counter pass is its own run (rocprofv3 collects one counter set per
run), written with ``--output-format csv``:

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir> -o run -- python3 bench.py --no-cpu --no-2d ...
    rocprofv3 --pmc WRITE_SIZE ...
    rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU ...
    rocprofv3 --pmc FETCH_SIZE ... -- python3 bench.py --calibrate 256   (and WRITE_SIZE)

``python -m phi_eng_amd.counters --fetch D1 --write D2 --valu D3
--calib-fetch D4 --calib-write D5 --copy-mib 256 --commit SHA`` then
writes the per-launch figures of each kernel kind (the names bench.py's
HIP-event timing uses).  FETCH_SIZE and WRITE_SIZE are scaled so that the
plain device copy of a known size reads and writes exactly its bytes.
"""
import argparse
import collections
import csv
import glob
import json
import os

# kernel function -> the kind name of the engine's timing table; a kind
# whose substep runs several launches (the Verlet density pass: rebuild,
# halo build and walk, those not needed returning at once) sums them.
KINDS = {
    "k_cell_key": "cell_key", "k_scan_reduce": "scan_reduce", "k_scan_partials": "scan_partials",
    "k_scan_apply": "scan_apply", "k_scan_apply_sum": "scan_apply", "k_scatter": "scatter", "k_rank_gather": "rank_gather",
    "k_density_list<false>": "density", "k_density_verlet": "density",
    "k_force_list": "force_integrate", "k_force_verlet": "force_integrate",
    "k_density_list<true>": "density_list_skin", "k_sur_step": "surrogate_step", "k_tile_vmax": "tile_vmax",
    "k_list_radius": "list_radius", "k_err_reduce": "err_reduce", "k_monitor": "monitor",
}
COPY_NAMES = ("__amd_rocclr_copyBuffer", "copyBuffer")


def _rows(d):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    with open(files[0]) as f:
        return list(csv.DictReader(f))


def _short(name):
    base = name.split("(")[0].strip()
    if base.startswith("void "):
        base = base[5:]
    return base


def per_dispatch(d):
    """{dispatch id: (kernel, grid, {counter: value})}"""
    out = {}
    for r in _rows(d):
        k = r.get("Dispatch_Id") or r.get("Correlation_Id")
        name = _short(r["Kernel_Name"])
        grid = int(float(r.get("Grid_Size", 0) or 0))
        ent = out.setdefault(k, (name, grid, collections.defaultdict(float)))
        ent[2][r["Counter_Name"]] += float(r["Counter_Value"])
    return out


def per_kind(d, counters):
    """Per kind: the mean over its launches of each counter, launches of
    one kind in a substep added up (SQ_WAVES: the largest of them, the
    grid of one pass)."""
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for name, grid, vals in per_dispatch(d).values():
        kind = KINDS.get(name)
        if kind is None:
            continue
        for c in counters:
            acc[kind][name].append(vals.get(c, 0.0))
    out = {}
    for kind, names in acc.items():
        out[kind] = {c: 0.0 for c in counters}
        for name, v in names.items():
            n = len(v) // len(counters)
            for i, c in enumerate(counters):
                mean = sum(v[i::len(counters)]) / max(n, 1)
                out[kind][c] = max(out[kind][c], mean) if c == "SQ_WAVES" else out[kind][c] + mean
    return out


def copy_kib(d, counter):
    vals = [v[2].get(counter, 0.0) for v in per_dispatch(d).values() if any(c in v[0] for c in COPY_NAMES)]
    if not vals:
        raise SystemExit(f"no copy dispatch under {d}")
    return sorted(vals)[len(vals) // 2]


def main():
    ap = argparse.ArgumentParser()
    for k in ("fetch", "write", "valu", "calib_fetch", "calib_write"):
        ap.add_argument("--" + k.replace("_", "-"), required=True)
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--commit", required=True)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                  "profiles"))
    a = ap.parse_args()
    kib = a.copy_mib * 1024.0
    fscale = kib / copy_kib(a.calib_fetch, "FETCH_SIZE")
    wscale = kib / copy_kib(a.calib_write, "WRITE_SIZE")
    fe = per_kind(a.fetch, ["FETCH_SIZE"])
    wr = per_kind(a.write, ["WRITE_SIZE"])
    traffic = {
        "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (separate runs), KiB per dispatch of the "
                  "1,048,576-particle launches, each scaled by what it reads for a "
                  f"{a.copy_mib} MiB device copy",
        "workload": "dam_break_3d_1m (1,048,576 particles)",
        "commit": a.commit,
        "fetch_scale": round(fscale, 4),
        "write_scale": round(wscale, 4),
        "bytes_per_launch": {k: int(round(1024.0 * (fe.get(k, {}).get("FETCH_SIZE", 0.0) * fscale +
                                                    wr.get(k, {}).get("WRITE_SIZE", 0.0) * wscale)))
                             for k in sorted(set(fe) | set(wr))},
    }
    va = per_kind(a.valu, ["SQ_INSTS_VALU", "SQ_WAVES"])
    valu = {
        "source": "rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU (per dispatch, 1,048,576-particle launches); "
                  "VALU issue peak = 1,024 SIMDs x one wave instruction per 4 cycles",
        "workload": "dam_break_3d_1m (1,048,576 particles)",
        "commit": a.commit,
        "per_launch": {k: {"insts_valu": int(round(v["SQ_INSTS_VALU"])), "waves": int(round(v["SQ_WAVES"])),
                           "insts_per_wave": round(v["SQ_INSTS_VALU"] / max(v["SQ_WAVES"], 1.0), 1)}
                       for k, v in sorted(va.items())},
    }
    with open(os.path.join(a.out, "traffic.json"), "w") as f:
        json.dump(traffic, f, indent=1)
        f.write("\n")
    with open(os.path.join(a.out, "valu.json"), "w") as f:
        json.dump(valu, f, indent=1)
        f.write("\n")
    print(json.dumps({"traffic": traffic["bytes_per_launch"], "valu": valu["per_launch"]}, indent=1))


if __name__ == "__main__":
    main()

"""Summaries of rocprofv3 output for the engine's kernels.

    python -m cluestering_alpaka3_amd.profiling stats <kernel_stats.csv>
        per-kernel call count and average duration (a --kernel-trace --stats run)

    python -m cluestering_alpaka3_amd.profiling run-batched [--events 4096] [--steps 5] [--ppb 64]
                                                            [--nh-first-radius 1.0]
        the batched calorimeter workload (the events of bench.py's batched
        measurement) clustered --steps times in one batched call each, with the
        given points per tile and tuning: the program to run under rocprofv3

    python -m cluestering_alpaka3_amd.profiling sq <pmc.csv> [--top 12]
        per-kernel SQ counters of a --pmc run (SQ_WAVES, SQ_INSTS_*, SQ_BUSY_CYCLES):
        instructions per wave, and the VALU instructions per second of the kernel
        (with its mean duration from a --kernel-trace --stats summary, --stats-csv)
        against one wave64 VALU instruction per 4 cycles on each of 1024 SIMDs at
        2.4 GHz

    python -m cluestering_alpaka3_amd.profiling occupancy <pmc.csv>
        per-kernel mean of rocprofv3's derived OccupancyPercent (resident waves as
        a share of 32 per CU over the kernel's active time) and
        MeanOccupancyPerActiveCU (waves per CU while the CU is busy)

    python -m cluestering_alpaka3_amd.profiling resources <remarks.txt> [--match k_density ...]
        per-kernel VGPRs, SGPRs, LDS bytes, scratch and the compiler's waves per
        SIMD, from hipcc -Rpass-analysis=kernel-resource-usage output

    python -m cluestering_alpaka3_amd.profiling traffic <fetch.csv> <write.csv> [--copy-mib 64] [--json OUT]
        HBM bytes per dispatch from two --pmc runs (FETCH_SIZE, then WRITE_SIZE),
        each scaled by what the same counter reads for the device-to-device copy
        of known size that bench.py makes before timing (--calibration-copy-mib):
        the copy moves copy_mib MiB in and copy_mib MiB out, so
        factor = copy bytes / counter bytes of that copy dispatch.
"""
from __future__ import annotations

import argparse
import csv
import json
import re
from collections import defaultdict

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short_name(kernel_name):
    """'void clue_hip::k_density<2, 0, 1, true, true>(clue_hip::PassArgs)' -> 'k_density'."""
    base = kernel_name.split("(")[0]
    base = re.sub(r"<.*>", "", base)
    return base.split("::")[-1].replace("void ", "").strip()


def kernel_stats(path):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((r["Name"], int(r["Calls"]), float(r["AverageNs"]), float(r["Percentage"])))
    rows.sort(key=lambda x: -x[3])
    return rows


def _per_dispatch(path, counter):
    vals = defaultdict(list)
    grids = defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            vals[r["Kernel_Name"]].append(float(r["Counter_Value"]))
            grids[r["Kernel_Name"]].append(int(r["Grid_Size"]))
    return vals, grids


def traffic(fetch_csv, write_csv, copy_mib=64):
    fetch, fgrid = _per_dispatch(fetch_csv, "FETCH_SIZE")
    write, wgrid = _per_dispatch(write_csv, "WRITE_SIZE")
    copy_kib = copy_mib * 1024.0

    def calib(vals, grids):
        # the calibration copy is the largest copy dispatch of the run
        cands = [(g, v) for name in vals if name.startswith(COPY_KERNEL)
                 for g, v in zip(grids[name], vals[name])]
        g, v = max(cands)
        return copy_kib / v, v

    ff, fraw = calib(fetch, fgrid)
    fw, wraw = calib(write, wgrid)
    out = {"calibration": {"copy_mib": copy_mib, "fetch_counter_kib": fraw, "write_counter_kib": wraw,
                           "fetch_factor": ff, "write_factor": fw},
           "kernels": {}}
    for name in fetch:
        if name.startswith(COPY_KERNEL) or name.startswith("__amd"):
            continue
        fk = sum(fetch[name]) / len(fetch[name])
        wk = sum(write.get(name, [0.0])) / max(1, len(write.get(name, [])))
        key = short_name(name)
        prev = out["kernels"].get(key)
        fb, wb = fk * 1024 * ff, wk * 1024 * fw
        if prev is None or fb + wb > prev["traffic_bytes"]:
            out["kernels"][key] = {"name": name, "dispatches": len(fetch[name]), "fetch_bytes": round(fb),
                                   "write_bytes": round(wb), "traffic_bytes": round(fb) + round(wb)}
    return out


VALU_PEAK = 1024 * 2.4e9 / 4  # wave64 VALU instructions per second (256 CUs x 4 SIMDs)


def sq_counters(path):
    """Mean of each SQ counter per dispatch, per kernel (short name)."""
    agg = defaultdict(lambda: defaultdict(list))
    for r in csv.DictReader(open(path)):
        agg[short_name(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in agg.items()}


def occupancy(path):
    """Mean of each derived occupancy counter per dispatch, per kernel (full template name)."""
    agg = defaultdict(lambda: defaultdict(list))
    for r in csv.DictReader(open(path)):
        agg[r["Kernel_Name"].split("(")[0].replace("void ", "")][r["Counter_Name"]].append(
            float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in agg.items()}


def resources(path, match):
    """Kernel resource usage from the compiler's kernel-resource-usage remarks."""
    rows = {}
    name = None
    for line in open(path):
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            rows[name] = {}
            continue
        m = re.search(r"remark:\s+([A-Za-z ]+?)(?: \[[^\]]*\])?: (\d+)", line)
        if name and m:
            rows[name][m.group(1).strip()] = int(m.group(2))
    keep = {k: v for k, v in rows.items() if not match or any(s in k for s in match)}
    return keep


def run_batched(events, steps, ppb, nh_first_radius, xcd_remap=None, pending_group=None, density_order=None,
                copy_mib=64):
    import numpy as np

    import cluestering_alpaka3_amd as clue
    from cluestering_alpaka3_amd import datasets

    sizes = datasets.event_sizes(events, 2024)
    evs = [datasets.calorimeter_event(int(n), 7_000_000 + k) for k, n in enumerate(sizes)]
    coords = np.ascontiguousarray(np.concatenate([e.coords for e in evs], axis=1))
    weights = np.ascontiguousarray(np.concatenate([e.weights for e in evs]))
    off = np.zeros(sizes.size + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    n = coords.shape[1]
    stream = clue.Stream(0)
    plan = clue.Plan(0)
    dcoords = clue.DeviceArray.from_host(coords)
    dw = clue.DeviceArray.from_host(weights)
    cp = [dcoords.offset(d * n) for d in range(2)]
    outs = [clue.DeviceArray((n,), t) for t in (np.int32, np.int32, np.float32, np.int32)]
    params = clue.make_params(1.3, 2.0, 2.6, ppb=ppb)
    tuning = plan.get_tuning()
    tuning["nh_first_radius"] = nh_first_radius
    for key, v in (("xcd_remap", xcd_remap), ("pending_group", pending_group), ("density_order", density_order)):
        if v is not None:
            tuning[key] = v
    plan.set_tuning(**tuning)
    if copy_mib > 0:  # the device-to-device copy of known size that calibrates FETCH_SIZE / WRITE_SIZE
        src = clue.DeviceArray(((copy_mib << 20) // 4,), np.float32)
        dst = clue.DeviceArray(((copy_mib << 20) // 4,), np.float32)
        clue.check(clue.lib().clue_memcpy_d2d(dst.ptr, src.ptr, copy_mib << 20, stream.handle))
        stream.synchronize()
        src.free()
        dst.free()
    kd, md = clue.kernel_desc(("flat", 0.5)), clue.metric_desc("euclidean")
    for step in range(steps):
        plan.make_clusters_batched_device(stream, params, 2, off, cp, dw.ptr, kd, md, None, *[o.ptr for o in outs])
        stream.synchronize()
        print(f"step {step + 1}/{steps}: {events} events, {n} points", flush=True)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("run-batched")
    b.add_argument("--events", type=int, default=4096)
    b.add_argument("--steps", type=int, default=5)
    b.add_argument("--ppb", type=int, default=64)
    b.add_argument("--nh-first-radius", type=float, default=1.0)
    b.add_argument("--xcd-remap", type=int)
    b.add_argument("--pending-group", type=int)
    b.add_argument("--density-order", type=int)
    b.add_argument("--copy-mib", type=int, default=64)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--top", type=int, default=25)
    q = sub.add_parser("sq")
    q.add_argument("csv")
    q.add_argument("--stats-csv")
    q.add_argument("--top", type=int, default=12)
    o = sub.add_parser("occupancy")
    o.add_argument("csv")
    r = sub.add_parser("resources")
    r.add_argument("remarks")
    r.add_argument("--match", nargs="*", default=[])
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--copy-mib", type=int, default=64)
    t.add_argument("--json")
    a = ap.parse_args()
    if a.cmd == "run-batched":
        run_batched(a.events, a.steps, a.ppb, a.nh_first_radius, a.xcd_remap, a.pending_group, a.density_order,
                    a.copy_mib)
    elif a.cmd == "sq":
        cnt = sq_counters(a.csv)
        dur = {}
        if a.stats_csv:
            for name, calls, avg, pct in kernel_stats(a.stats_csv):
                dur.setdefault(short_name(name), avg * 1e-9)
        print(f"{'kernel':26s} {'waves':>8s} {'valu/w':>9s} {'salu/w':>8s} {'smem/w':>7s} {'lds/w':>7s} "
              f"{'vmem/w':>7s} {'valu_instr/s':>13s} {'of_peak':>8s}")
        rows = sorted(cnt.items(), key=lambda kv: -kv[1].get("SQ_INSTS_VALU", 0))
        for k, c in rows[: a.top]:
            w = max(c.get("SQ_WAVES", 1.0), 1.0)
            vm = c.get("SQ_INSTS_VMEM_RD", 0.0) + c.get("SQ_INSTS_VMEM_WR", 0.0)
            line = (f"{k:26s} {w:8.0f} {c.get('SQ_INSTS_VALU', 0) / w:9.1f} {c.get('SQ_INSTS_SALU', 0) / w:8.1f} "
                    f"{c.get('SQ_INSTS_SMEM', 0) / w:7.1f} {c.get('SQ_INSTS_LDS', 0) / w:7.1f} {vm / w:7.1f}")
            if k in dur:
                rate = c.get("SQ_INSTS_VALU", 0) / dur[k]
                line += f" {rate:13.3e} {rate / VALU_PEAK:8.3f}"
            print(line)
    elif a.cmd == "occupancy":
        res = occupancy(a.csv)
        print(f"{'kernel':64s} {'OccupancyPercent':>17s} {'waves/active CU':>16s}")
        for k, c in sorted(res.items(), key=lambda kv: kv[0]):
            print(f"{k[:64]:64s} {c.get('OccupancyPercent', float('nan')):17.1f} "
                  f"{c.get('MeanOccupancyPerActiveCU', float('nan')):16.2f}")
    elif a.cmd == "resources":
        res = resources(a.remarks, a.match)
        print(f"{'kernel (mangled)':72s} {'VGPRs':>6s} {'SGPRs':>6s} {'LDS_B':>6s} {'scratch':>8s} {'waves/SIMD':>10s}")
        for k, v in res.items():
            print(f"{k[:72]:72s} {v.get('VGPRs', -1):6d} {v.get('TotalSGPRs', -1):6d} "
                  f"{v.get('LDS Size', -1):6d} {v.get('ScratchSize', -1):8d} {v.get('Occupancy', -1):10d}")
    elif a.cmd == "stats":
        print(f"{'kernel':100s} {'calls':>6s} {'avg_us':>10s} {'share%':>7s}")
        for name, calls, avg, pct in kernel_stats(a.csv)[: a.top]:
            print(f"{name[:100]:100s} {calls:6d} {avg / 1e3:10.2f} {pct:7.2f}")
    else:
        res = traffic(a.fetch_csv, a.write_csv, a.copy_mib)
        c = res["calibration"]
        print(f"calibration: {c['copy_mib']} MiB copy read as FETCH {c['fetch_counter_kib']:.1f} KiB "
              f"(x{c['fetch_factor']:.4f}), WRITE {c['write_counter_kib']:.1f} KiB (x{c['write_factor']:.4f})")
        print(f"{'kernel':28s} {'dispatches':>10s} {'fetch_MB':>10s} {'write_MB':>10s} {'total_MB':>10s}")
        for k, v in sorted(res["kernels"].items(), key=lambda kv: -kv[1]["traffic_bytes"]):
            print(f"{k:28s} {v['dispatches']:10d} {v['fetch_bytes'] / 1e6:10.2f} {v['write_bytes'] / 1e6:10.2f} "
                  f"{v['traffic_bytes'] / 1e6:10.2f}")
        if a.json:
            with open(a.json, "w") as f:
                json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()

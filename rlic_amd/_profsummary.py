"""Summaries of rocprofv3 CSV output for the LIC kernels.

    python -m rlic_amd._profsummary stats  <kernel_stats.csv> [--templates]
    python -m rlic_amd._profsummary traffic <FETCH_SIZE counter_collection.csv>
                                            <WRITE_SIZE counter_collection.csv>
                                            [--workload NAME] [--out-dir DIR]
    python -m rlic_amd._profsummary counters <counter_collection.csv> ... [--templates]
    python -m rlic_amd._profsummary valu <counter_collection.csv> ...
                                         [--workload NAME] [--out-dir DIR]

`stats` prints the per-kernel launch count and average duration.
`traffic` turns the per-dispatch FETCH_SIZE / WRITE_SIZE counters (KiB)
into HBM bytes per launch.  Each counter is first calibrated on the
library's copy kernel (lic_copy2d, run by ``bench.py --calibrate-copy`` on
2^28 float32 = 1 GiB read and 1 GiB written per dispatch): factor = known
bytes / counter bytes.  With --out-dir, writes traffic_<kernel>_<workload>.json
files that bench.py reports as roofline.traffic.
`counters` prints, per kernel, the ratios of the counters present in the
given runs: VALU lane utilization SQ_THREAD_CYCLES_VALU /
(64 * SQ_ACTIVE_INST_VALU), LDS bank-conflict share SQ_LDS_BANK_CONFLICT /
SQ_LDS_IDX_ACTIVE, L2 hit rate TCC_HIT / (TCC_HIT + TCC_MISS), L1 hit rate
1 - TCP_TCC_READ_REQ / TCP_TOTAL_ACCESSES, VALU / SALU / LDS / vector-read
instructions per wave, and the share of wave cycles spent waiting
SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES.
`valu` reads SQ_INSTS_VALU and SQ_WAVES (and, when present, GRBM_GUI_ACTIVE,
SQ_BUSY_CYCLES, SQ_ACTIVE_INST_VALU) per dispatch and, with --out-dir, writes
valu_<kernel>_<workload>.json with the wave64 VALU instructions per launch,
which bench.py turns into the VALU issue rate of its roofline.
With --templates, kernels are told apart by their template arguments
(e.g. the tile shapes of lic_replay_lds) instead of by name alone.
"""

from __future__ import annotations

import argparse
import csv
import json
import re
import sys
from collections import defaultdict
from pathlib import Path

COPY_KERNEL = "lic_copy2d"
COPY_BYTES = 2**28 * 4  # one lic_copy2d dispatch of the calibration array, each way


def short_name(full: str, templates: bool = False) -> str:
    """'void rlic::lic_trace<float, ...>(...)' -> 'lic_trace' (or, with
    templates, 'lic_trace<float, ...>')."""
    m = re.search(r"rlic::(\w+<[^()]*>)" if templates else r"rlic::(\w+)", full)
    if m is None and templates:
        m = re.search(r"rlic::(\w+)", full)
    return m.group(1) if m else full.split("(")[0][:60]


def read_stats(path: Path, templates: bool = False) -> dict[str, dict]:
    out: dict[str, dict] = {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            name = short_name(row["Name"], templates)
            agg = out.setdefault(name, {"calls": 0, "total_ns": 0.0})
            agg["calls"] += int(row["Calls"])
            agg["total_ns"] += float(row["TotalDurationNs"])
    for agg in out.values():
        agg["avg_ms"] = agg["total_ns"] / agg["calls"] / 1e6
    return out


def read_counter(path: Path, counter: str) -> dict[str, list[float]]:
    """Per kernel: the counter's value (summed over instances) per dispatch."""
    per_dispatch: dict[tuple[str, str], float] = defaultdict(float)
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] != counter:
                continue
            key = (short_name(row["Kernel_Name"]), row["Dispatch_Id"])
            per_dispatch[key] += float(row["Counter_Value"])
    out: dict[str, list[float]] = defaultdict(list)
    for (name, _), value in per_dispatch.items():
        out[name].append(value)
    return out


def traffic(fetch_csv: Path, write_csv: Path) -> dict[str, dict]:
    fetch = read_counter(fetch_csv, "FETCH_SIZE")
    write = read_counter(write_csv, "WRITE_SIZE")
    if COPY_KERNEL not in fetch or COPY_KERNEL not in write:
        raise SystemExit(f"no {COPY_KERNEL} dispatches: run bench.py with --calibrate-copy")
    avg = lambda xs: sum(xs) / len(xs)  # noqa: E731
    f_cal = COPY_BYTES / (avg(fetch[COPY_KERNEL]) * 1024)
    w_cal = COPY_BYTES / (avg(write[COPY_KERNEL]) * 1024)
    result = {}
    for name in sorted(set(fetch) & set(write)):
        if not name.startswith("lic_"):
            continue
        fb = avg(fetch[name]) * 1024 * f_cal
        wb = avg(write[name]) * 1024 * w_cal
        result[name] = {
            "fetch_bytes_per_launch": fb,
            "write_bytes_per_launch": wb,
            "hbm_bytes_per_launch": fb + wb,
            "fetch_kib_raw": avg(fetch[name]),
            "write_kib_raw": avg(write[name]),
            "dispatches": len(fetch[name]),
            "fetch_calibration": f_cal,
            "write_calibration": w_cal,
        }
    return result


def counter_ratios(paths: list[Path], templates: bool = False) -> dict[str, dict]:
    """Per kernel: counter totals over all dispatches and derived ratios."""
    totals: dict[str, dict[str, float]] = defaultdict(lambda: defaultdict(float))
    dispatches: dict[str, set] = defaultdict(set)
    for path in paths:
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                name = short_name(row["Kernel_Name"], templates)
                counter = row["Counter_Name"].removesuffix("_sum")
                totals[name][counter] += float(row["Counter_Value"])
                dispatches[name].add((str(path), row["Dispatch_Id"]))
    out = {}
    for name, c in totals.items():
        r: dict[str, float] = {}
        if c.get("SQ_ACTIVE_INST_VALU"):
            r["valu_lane_utilization"] = c["SQ_THREAD_CYCLES_VALU"] / (64 * c["SQ_ACTIVE_INST_VALU"])
        if c.get("SQ_LDS_IDX_ACTIVE"):
            r["lds_bank_conflict_share"] = c["SQ_LDS_BANK_CONFLICT"] / c["SQ_LDS_IDX_ACTIVE"]
        if c.get("TCC_HIT", 0) + c.get("TCC_MISS", 0):
            r["l2_hit_rate"] = c["TCC_HIT"] / (c["TCC_HIT"] + c["TCC_MISS"])
        if c.get("TCP_TOTAL_ACCESSES"):
            r["l1_hit_rate"] = 1 - c["TCP_TCC_READ_REQ"] / c["TCP_TOTAL_ACCESSES"]
        if c.get("SQ_WAVES"):
            for k in ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD"):
                if k in c:
                    r[k.lower().replace("sq_insts_", "") + "_per_wave"] = c[k] / c["SQ_WAVES"]
        if c.get("SQ_WAVE_CYCLES") and "SQ_WAIT_INST_ANY" in c:
            r["wait_share"] = c["SQ_WAIT_INST_ANY"] / c["SQ_WAVE_CYCLES"]
        out[name] = {"ratios": r, "totals": dict(c), "dispatches": len(dispatches[name])}
    return out


def valu(paths: list[Path]) -> dict[str, dict]:
    """Per kernel: wave64 VALU instructions and waves per launch (averages
    over its dispatches), with the other counters present, per launch."""
    per: dict[str, dict[str, dict[str, float]]] = defaultdict(lambda: defaultdict(dict))
    for path in paths:
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                name = short_name(row["Kernel_Name"])
                counter = row["Counter_Name"].removesuffix("_sum")
                key = f"{path}:{row['Dispatch_Id']}"
                d = per[name][key]
                d[counter] = d.get(counter, 0.0) + float(row["Counter_Value"])
    out = {}
    for name, disp in per.items():
        rows = [d for d in disp.values() if "SQ_INSTS_VALU" in d and "SQ_WAVES" in d]
        if not rows:
            continue
        avg = {c: sum(d.get(c, 0.0) for d in rows) / len(rows)
               for c in sorted({c for d in rows for c in d})}
        out[name] = {"valu_insts_per_launch": avg["SQ_INSTS_VALU"],
                     "waves_per_launch": avg["SQ_WAVES"],
                     "valu_insts_per_wave": avg["SQ_INSTS_VALU"] / avg["SQ_WAVES"],
                     "counters_per_launch": avg, "dispatches": len(rows)}
    return out


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv", type=Path)
    s.add_argument("--templates", action="store_true")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv", type=Path)
    t.add_argument("write_csv", type=Path)
    t.add_argument("--workload", default="plain4096")
    t.add_argument("--out-dir", type=Path)
    c = sub.add_parser("counters")
    c.add_argument("csv", type=Path, nargs="+")
    c.add_argument("--templates", action="store_true")
    v = sub.add_parser("valu")
    v.add_argument("csv", type=Path, nargs="+")
    v.add_argument("--workload", default="plain4096")
    v.add_argument("--out-dir", type=Path)
    args = ap.parse_args(argv)
    if args.cmd == "valu":
        for name, r in sorted(valu(args.csv).items()):
            if not name.startswith("lic_"):
                continue
            print(f"{name:24s} dispatches {r['dispatches']}, VALU/launch "
                  f"{r['valu_insts_per_launch']:.4g}, waves {r['waves_per_launch']:.0f}, "
                  f"VALU/wave {r['valu_insts_per_wave']:.1f}")
            if args.out_dir:
                path = args.out_dir / f"valu_{name}_{args.workload}.json"
                path.write_text(json.dumps(r, indent=1) + "\n")
        return 0
    if args.cmd == "counters":
        for name, r in sorted(counter_ratios(args.csv, args.templates).items()):
            if not name.startswith("lic_"):
                continue
            ratios = ", ".join(f"{k} {v:.4g}" for k, v in r["ratios"].items())
            print(f"{name:24s} dispatches {r['dispatches']}, {ratios}")
        return 0
    if args.cmd == "stats":
        for name, agg in sorted(read_stats(args.csv, args.templates).items(),
                                key=lambda kv: -kv[1]["total_ns"]):
            print(f"{name:24s} calls {agg['calls']:6d}  avg {agg['avg_ms']:10.4f} ms  "
                  f"total {agg['total_ns'] / 1e6:10.3f} ms")
        return 0
    res = traffic(args.fetch_csv, args.write_csv)
    for name, r in res.items():
        print(f"{name:24s} HBM {r['hbm_bytes_per_launch'] / 1e9:9.3f} GB/launch "
              f"(fetch {r['fetch_bytes_per_launch'] / 1e9:.3f}, write "
              f"{r['write_bytes_per_launch'] / 1e9:.3f}; calibration x{r['fetch_calibration']:.3f}"
              f" / x{r['write_calibration']:.3f})")
        if args.out_dir:
            path = args.out_dir / f"traffic_{name}_{args.workload}.json"
            path.write_text(json.dumps(r, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Hardware-counter summaries of the move kernels, from rocprofv3 runs.

Three commands:

  python3 -m wacommplusplus_amd.counters copy [--mib 1024] [--reps 5]
      A plain device-to-device copy of known size, run under rocprofv3 with
      FETCH_SIZE and then WRITE_SIZE to scale those counters to bytes.

  python3 -m wacommplusplus_amd.counters parse --runs RUNS.json --sq DIR
        --fetch DIR --write DIR --copy-fetch DIR --copy-write DIR
        [--trace DIR] [--out profiles/move_counters.json] [--md FILE]
      Reads the counter CSVs of `bench.py --sort-tiles ...` (whose JSON line,
      RUNS.json, lists the runs in launch order) and writes one entry per run
      key ("f64@t0", "bf16_tile_euler@t1", ...) holding the per-interval
      values of the run's timed launches: VALU, VMEM-read and SALU wave
      instructions, TCC hits and misses, and the calibrated L2-to-memory
      bytes.  The move launches of one variant share one kernel symbol, so
      each symbol's dispatches are assigned to that variant's runs in launch
      order (warm-up launches skipped).

  python3 -m wacommplusplus_amd.counters bench-trace --bench OUT.json
        --moves MOVES.csv [--stats STATS.csv] --md FILE
      Splits the move launches of a default `bench.py` run's kernel trace
      into the bench's runs and sets their mean durations beside the
      HIP-event means of the bench line.
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import sys
from collections import OrderedDict, defaultdict

MOVE_KERNELS = ("move_parity_kernel", "move_fast_kernel", "move_fast_lds_kernel")


def copy_probe(mib: int, reps: int) -> None:
    import torch
    n = mib * (1 << 20) // 4
    a = torch.arange(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    assert torch.equal(a[-8:], b[-8:])
    print(json.dumps({"copy_bytes": n * 4, "reps": reps}), flush=True)


def read_counters(d: str) -> "OrderedDict[int, dict]":
    """Dispatch id -> {"name", counters...} from one rocprofv3 counter run."""
    path = os.path.join(d, "run_counter_collection.csv")
    out: "OrderedDict[int, dict]" = OrderedDict()
    with open(path) as f:
        for r in csv.DictReader(f):
            k = int(r["Dispatch_Id"])
            e = out.setdefault(k, {"name": r["Kernel_Name"], "grid": int(r["Grid_Size"])})
            e[r["Counter_Name"]] = e.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return OrderedDict(sorted(out.items()))


def short_name(name: str) -> str:
    """A readable kernel name; rocprofv3 leaves symbols with _Float16
    template arguments mangled, which are read here as far as needed."""
    if name.startswith("_ZN"):
        rest = name.split("_GLOBAL__N_1", 1)[-1]
        digits = len(rest) - len(rest.lstrip("0123456789"))
        base = rest[digits:digits + int(rest[:digits])] if digits else name
        if "_kernel" not in base:
            return name
        tail = rest[digits + len(base):]
        args = ["f16" if "DF16_" in tail else "?"]
        if "TiledPacked" in tail:
            args.append("TiledPacked<f16>")
        elif "Packed" in tail:
            args.append("Packed<f16>")
        return f"{base}<{', '.join(args)}>"
    for k in MOVE_KERNELS:
        if k in name:
            return k + name.split(k, 1)[1].split(">(", 1)[0].replace("wc::(anonymous namespace)::", "") + ">"
    return name.replace("wc::(anonymous namespace)::", "").split("(", 1)[0].replace("void ", "")


def move_dispatches(disp: "OrderedDict[int, dict]") -> "OrderedDict[str, list]":
    """Move-kernel dispatches grouped by symbol, symbols in first-launch order."""
    out: "OrderedDict[str, list]" = OrderedDict()
    for k, e in disp.items():
        if any(m in e["name"] for m in MOVE_KERNELS):
            out.setdefault(e["name"], []).append(e)
    return out


def copy_scale(d: str, counter: str, copy_bytes: int) -> dict:
    """Bytes per counter unit: the copy's known bytes over what the counter
    read for the longest-running dispatch of the probe (the copy itself)."""
    disp = read_counters(d)
    vals = [e[counter] for e in disp.values() if counter in e]
    top = max(vals)
    copies = [v for v in vals if v > 0.5 * top]
    mean = sum(copies) / len(copies)
    return {"counter": counter, "copy_bytes": copy_bytes, "reading": mean, "dispatches": len(copies),
            "bytes_per_unit": copy_bytes / mean}


def assign(runs: list, groups: "OrderedDict[str, list]") -> dict:
    """Run key -> (kernel symbol, the run's timed dispatches)."""
    variants = list(OrderedDict((r["variant"], None) for r in runs))
    if len(variants) != len(groups):
        raise SystemExit(f"{len(variants)} variants in the runs but {len(groups)} move-kernel symbols")
    out = {}
    for v, (name, ds) in zip(variants, groups.items()):
        pos = 0
        for r in (r for r in runs if r["variant"] == v):
            pos += r["warmup_launches"]
            out[r["key"]] = (name, ds[pos:pos + r["timed_launches"]])
            pos += r["timed_launches"]
        if pos != len(ds):
            raise SystemExit(f"{v}: {len(ds)} dispatches of {short_name(name)}, the runs account for {pos}")
    return out


def parse(a) -> None:
    with open(a.runs) as f:
        line = [json.loads(x) for x in f if x.startswith("{")][-1]
    runs = line["profile_runs"]
    copy_bytes = a.copy_bytes
    cal_f = copy_scale(a.copy_fetch, "FETCH_SIZE", copy_bytes)
    cal_w = copy_scale(a.copy_write, "WRITE_SIZE", copy_bytes)
    sq = assign(runs, move_dispatches(read_counters(a.sq)))
    fe = assign(runs, move_dispatches(read_counters(a.fetch)))
    wr = assign(runs, move_dispatches(read_counters(a.write)))
    kernels = OrderedDict()
    for r in runs:
        name, ds = sq[r["key"]]
        fd, wd = fe[r["key"]][1], wr[r["key"]][1]
        fetch = [e["FETCH_SIZE"] * cal_f["bytes_per_unit"] for e in fd]
        write = [e["WRITE_SIZE"] * cal_w["bytes_per_unit"] for e in wd]
        hit = sum(e["TCC_HIT"] for e in ds)
        miss = sum(e["TCC_MISS"] for e in ds)
        kernels[r["key"]] = {
            "kernel": short_name(name),
            "sort_tile": r["sort_tile"],
            "bench_launch_ms": r["move_kernel_ms"],
            "waves_per_dispatch": ds[0]["SQ_WAVES"],
            "valu_per_interval": [e["SQ_INSTS_VALU"] for e in ds],
            "vmem_rd_per_interval": [e["SQ_INSTS_VMEM_RD"] for e in ds],
            "salu_per_interval": [e["SQ_INSTS_SALU"] for e in ds],
            "tcc_hit_rate": hit / max(1.0, hit + miss),
            "fetch_bytes_per_interval": fetch,
            "write_bytes_per_interval": write,
            "mem_bytes_per_interval": [x + y for x, y in zip(fetch, write)],
        }
    # every other kernel of the counter runs: mean per dispatch
    others = defaultdict(lambda: defaultdict(list))
    for d, cs in ((a.sq, ("SQ_INSTS_VALU", "SQ_WAVES")), (a.fetch, ("FETCH_SIZE",)), (a.write, ("WRITE_SIZE",))):
        for e in read_counters(d).values():
            if any(m in e["name"] for m in MOVE_KERNELS):
                continue
            for c in cs:
                others[short_name(e["name"])][c].append(e[c])
    other = {}
    for n, cs in sorted(others.items()):
        o = {c: sum(v) / len(v) for c, v in cs.items()}
        o["dispatches"] = len(cs["SQ_WAVES"])
        if "FETCH_SIZE" in o:
            o["fetch_bytes"] = o.pop("FETCH_SIZE") * cal_f["bytes_per_unit"]
        if "WRITE_SIZE" in o:
            o["write_bytes"] = o.pop("WRITE_SIZE") * cal_w["bytes_per_unit"]
        other[n] = o
    out = {
        "what": "per-interval hardware counters of the move kernels (intervals 0 .. n-1 from t0 of the bench "
                "workload), one entry per run key variant[_euler]@t<cell-tile sort>",
        "commands": a.command or [],
        "steps": line["steps"], "warmup": line["warmup"], "particles": line["particles"],
        "calibration": {"fetch": cal_f, "write": cal_w,
                        "note": "FETCH_SIZE and WRITE_SIZE scaled by their reading for a plain device copy of "
                                "known size; they count L2-to-memory traffic, which the Infinity Cache may "
                                "serve before HBM"},
        "kernels": kernels,
        "other_kernels_mean_per_dispatch": other,
    }
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    if a.md:
        write_md(a, out, runs)


def trace_by_run(d: str, runs: list) -> dict:
    """Run key -> mean kernel-trace duration (ms) of its timed move launches."""
    path = os.path.join(d, "run_kernel_trace.csv")
    disp = OrderedDict()
    with open(path) as f:
        for r in csv.DictReader(f):
            disp[int(r["Dispatch_Id"])] = {"name": r["Kernel_Name"],
                                           "ns": int(r["End_Timestamp"]) - int(r["Start_Timestamp"])}
    got = assign(runs, move_dispatches(OrderedDict(sorted(disp.items()))))
    return {k: sum(e["ns"] for e in ds) / len(ds) / 1e6 for k, (_, ds) in got.items()}


def write_md(a, out, runs) -> None:
    peak = 256 * 4 * 2.4e9 / 4
    rows = []
    trace = trace_by_run(a.trace, runs) if a.trace else {}
    for r in runs:
        k = out["kernels"][r["key"]]
        n = len(k["valu_per_interval"])
        valu = sum(k["valu_per_interval"]) / n
        mem = sum(k["mem_bytes_per_interval"]) / n
        ms = r["move_kernel_ms"]
        rows.append(f"| {r['key']} | {k['kernel']} | {ms:.3f} | {trace.get(r['key'], float('nan')):.3f} | "
                    f"{valu:.4g} | {valu / (ms / 1e3) / peak:.3f} | {mem / 1e6:.1f} | "
                    f"{mem / (ms / 1e3) / 1e9:.0f} | {k['tcc_hit_rate']:.3f} |")
    cal = out["calibration"]
    lines = [
        "# Move-kernel counters, full-size workload (4M particles, 1024x768x11, 24 intervals from t0)",
        "",
        "Produced by:",
        "",
        "```",
        *out["commands"],
        "```",
        "",
        f"Calibration (plain device copy of {cal['fetch']['copy_bytes'] / 2**20:.0f} MiB): FETCH_SIZE x "
        f"{cal['fetch']['bytes_per_unit']:.1f} B/unit, WRITE_SIZE x {cal['write']['bytes_per_unit']:.1f} B/unit.",
        "",
        "Per run (timed launches, intervals 0-23): launch time from the HIP events of the counter run's own "
        "bench line (`bench ms`) and from the kernel trace (`trace ms`, when given); VALU wave instructions per "
        "dispatch and their issue fraction against 256 CU x 4 SIMD x 2.4 GHz / 4 cycles at the bench launch "
        "time; calibrated FETCH+WRITE bytes per dispatch and their rate; L2 hit rate (TCC_HIT/(HIT+MISS)).",
        "",
        "| run | kernel | bench ms | trace ms | VALU/dispatch | VALU frac | L2-to-memory MB | GB/s | L2 hit |",
        "|---|---|---|---|---|---|---|---|---|",
        *rows,
        "",
        "Other kernels (mean per dispatch over the counter runs):",
        "",
        "| kernel | dispatches | waves | VALU | fetch MB | write MB |",
        "|---|---|---|---|---|---|",
    ]
    for n, o in out["other_kernels_mean_per_dispatch"].items():
        lines.append(f"| {n} | {o['dispatches']} | {o.get('SQ_WAVES', 0):.0f} | {o.get('SQ_INSTS_VALU', 0):.4g} | "
                     f"{o.get('fetch_bytes', 0) / 1e6:.2f} | {o.get('write_bytes', 0) / 1e6:.2f} |")
    if a.trace:
        lines += ["", "Kernel-trace statistics of the first command (rocprofv3 `--stats`, every dispatch of the run: "
                      "one symbol covers a variant's unsorted, sorted and Euler launches, split per run above):", "",
                  "| kernel | calls | total ms | mean us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
        with open(os.path.join(a.trace, "run_kernel_stats.csv")) as f:
            for r in list(csv.DictReader(f))[:30]:
                lines.append(f"| {short_name(r['Name'])} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.1f} | "
                             f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
                             f"{float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
    with open(a.md, "w") as f:
        f.write("\n".join(lines) + "\n")


AUTOTUNE_LAUNCHES = 6          # wc_autotune_sort: one warm-up trial, then tiles 0, 1, 2, 4, 8
TUNE_HOURS = 12                # bench.tune_sort moves the population to 12 h before tuning


def bench_trace(a) -> None:
    """The move launches of a default bench run (its kernel trace filtered to
    the move kernels) split into the bench's runs, and each run's mean trace
    duration set beside the HIP-event mean the bench line reports."""
    with open(a.bench) as f:
        line = [json.loads(x) for x in f if x.startswith("{")][-1]
    steps, warmup = line["steps"], line["warmup"]
    v = line["variants"]
    runs = []
    bases = [k for k in v if not k.endswith(("_unsorted", "_euler", "_lds"))]       # the variants, in run order
    for key in ["f64"] + bases:
        tile = line["config"]["cell_tile_sort"] if key == "f64" else v[key]["sort_tile"]
        runs.append(dict(key=key + "_unsorted", variant=key, warmup_launches=warmup, timed_launches=steps))
        if line["config"].get("sort_autotuned", True):
            runs.append(dict(key=key + "_tuning", variant=key, warmup_launches=TUNE_HOURS,
                             timed_launches=AUTOTUNE_LAUNCHES))
        if tile:
            runs.append(dict(key=key, variant=key, warmup_launches=warmup, timed_launches=steps))
        if key + "_euler" in v:
            runs.append(dict(key=key + "_euler", variant=key, warmup_launches=warmup, timed_launches=steps))
        if key + "_lds" in v:            # the LDS-staged move: a kernel symbol of its own
            runs.append(dict(key=key + "_lds", variant=key + "_lds", warmup_launches=warmup, timed_launches=steps))
    disp = OrderedDict()
    with open(a.moves) as f:
        rd = csv.reader(f)
        for r in rd:
            if r[0] == "Kind":
                continue
            disp[int(r[5])] = {"name": r[7], "ns": int(r[10]) - int(r[9])}
    groups = move_dispatches(OrderedDict(sorted(disp.items())))
    # launches after the variants' runs (precision, upload-inclusive rate, CPU-baseline snapshots) are
    # left over at the end of the parity symbol's list and ignored here
    variants = list(OrderedDict((r["variant"], None) for r in runs))
    rows = []
    for vk, (name, ds) in zip(variants, groups.items()):
        pos = 0
        for r in (r for r in runs if r["variant"] == vk):
            pos += r["warmup_launches"]
            seg = ds[pos:pos + r["timed_launches"]]
            pos += r["timed_launches"]
            if r["key"].endswith("_tuning"):
                continue
            bench_ms = (line["roofline"]["avg_launch_ms"] if r["key"] == "f64" else
                        v[r["key"]]["move_kernel_ms"])
            trace_ms = sum(e["ns"] for e in seg) / len(seg) / 1e6
            rows.append(f"| {r['key']} | {short_name(name)} | {len(seg)} | {bench_ms:.3f} | {trace_ms:.3f} | "
                        f"{100 * (trace_ms / bench_ms - 1):+.1f} % |")
    roof = line["roofline"]
    lines = [f"# Move kernels of the default benchmark: trace vs bench timing ({line['steps']} intervals, "
             f"{line['warmup']} warm-up)", "", "Produced by:", "", "```", *(a.command or []), "```", "",
             f"Bench line: value {line['value']:.4g} particle-steps/s (FP64 parity path, cell-tile sort "
             f"{line['config']['cell_tile_sort']}); roofline frac {roof['frac']} "
             f"({roof['counters']['key'] if roof.get('counters') else 'no counters'}).", "",
             "Each run's timed move launches, found in the kernel trace by launch order, against the mean "
             "the bench measured with HIP events on the context stream:", "",
             "| run | kernel | launches | bench ms | trace ms | trace - bench |", "|---|---|---|---|---|---|",
             *rows]
    if a.stats:
        lines += ["", "rocprofv3 `--stats` of the same run (every dispatch; one symbol covers several runs):", "",
                  "| kernel | calls | total ms | mean us | % |", "|---|---|---|---|---|"]
        with open(a.stats) as f:
            for r in list(csv.DictReader(f))[:40]:
                lines.append(f"| {short_name(r['Name'])} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.1f} | "
                             f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
    with open(a.md, "w") as f:
        f.write("\n".join(lines) + "\n")


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(prog="python3 -m wacommplusplus_amd.counters")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    c.add_argument("--reps", type=int, default=5)
    p = sub.add_parser("parse")
    for k in ("runs", "sq", "fetch", "write", "copy_fetch", "copy_write"):
        p.add_argument("--" + k.replace("_", "-"), required=True)
    p.add_argument("--copy-bytes", type=int, default=1024 << 20)
    p.add_argument("--trace", default=None)
    p.add_argument("--out", default="profiles/move_counters.json")
    p.add_argument("--md", default=None)
    p.add_argument("--command", action="append", help="a command the counters came from (repeatable)")
    b = sub.add_parser("bench-trace")
    b.add_argument("--bench", required=True, help="the bench run's stdout (its JSON line)")
    b.add_argument("--moves", required=True, help="its kernel-trace CSV, header and move-kernel rows only")
    b.add_argument("--stats", default=None, help="its rocprofv3 kernel_stats CSV")
    b.add_argument("--md", required=True)
    b.add_argument("--command", action="append")
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        copy_probe(a.mib, a.reps)
    elif a.cmd == "bench-trace":
        bench_trace(a)
    else:
        parse(a)


if __name__ == "__main__":
    main(sys.argv[1:])

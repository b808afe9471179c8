"""The counter tool (wacommplusplus_amd/counters.py) that turns rocprofv3 runs
into profiles/move_counters.json and the trace-vs-bench summary: its
assignment of move-kernel dispatches to the bench's runs, on synthetic
traces whose durations encode which run each dispatch belongs to, so a
shifted or mis-ordered assignment shows as a wrong mean."""
import csv
import json

from wacommplusplus_amd import counters as C

HEADER = ["Kind", "Agent_Id", "Queue_Id", "Stream_Id", "Thread_Id", "Dispatch_Id", "Kernel_Id", "Kernel_Name",
          "Correlation_Id", "Start_Timestamp", "End_Timestamp"]
NAMES = {"f64": "void wc::(anonymous namespace)::move_parity_kernel<wc::(anonymous namespace)::PackedSampler, "
                "wc::SoAAccess, false>(wc::MoveParams)",
         "i16_tile": "void wc::(anonymous namespace)::move_fast_kernel<short, wc::(anonymous namespace)::TiledI16, "
                     "false, false>(wc::MoveParams)",
         "bf16": "void wc::(anonymous namespace)::move_fast_kernel<unsigned short, wc::(anonymous namespace)::"
                 "Packed<unsigned short>, false, false>(wc::MoveParams)",
         "bf16_lds": "void wc::(anonymous namespace)::move_fast_lds_kernel<unsigned short>(wc::MoveParams)"}


def write_trace(path, seq):
    """seq: (variant, duration ns) in launch order, with other kernels between."""
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(HEADER)
        t, d = 1000, 1
        for v, ns in seq:
            w.writerow(["KERNEL_DISPATCH", "Agent 1", 0, 0, 0, d, 1, "wc::(anonymous namespace)::fill_u32_kernel()",
                        d, t, t + 5])
            d += 1
            w.writerow(["KERNEL_DISPATCH", "Agent 1", 0, 0, 0, d, 2, NAMES[v], d, t + 10, t + 10 + ns])
            d += 1
            t += 10 + ns + 100


def test_bench_trace_assigns_runs(tmp_path):
    steps, warmup = 3, 2
    # the bench's order per variant: unsorted (warmup + steps), tuning (12 + 6),
    # sorted if a tile was chosen (warmup + steps), Euler where reported
    plan = [("f64", "f64_unsorted", True), ("f64", "tune", False), ("f64", "f64", True), ("f64", "f64_euler", True),
            ("i16_tile", "i16_tile_unsorted", True), ("i16_tile", "tune", False),
            ("i16_tile", "i16_tile_euler", True),
            ("bf16", "bf16_unsorted", True), ("bf16", "tune", False), ("bf16", "bf16", True),
            ("bf16_lds", "bf16_lds", True)]
    ms = {"f64_unsorted": 25.0, "f64": 24.0, "f64_euler": 19.0, "i16_tile_unsorted": 13.0, "i16_tile_euler": 8.0,
          "bf16_unsorted": 11.0, "bf16": 9.0, "bf16_lds": 16.0}
    seq = []
    for v, key, timed in plan:
        if key == "tune":
            seq += [(v, 999_000)] * (C.TUNE_HOURS + C.AUTOTUNE_LAUNCHES)
        else:
            seq += [(v, 888_000)] * warmup + [(v, int(ms[key] * 1e6))] * steps
    seq += [("f64", 777_000)] * 5                       # later parity launches (precision, CPU snapshots)
    write_trace(tmp_path / "moves.csv", seq)
    variants = {
        "f64_unsorted": dict(move_kernel_ms=25.0, sort_tile=0), "f64_euler": dict(move_kernel_ms=19.0, sort_tile=1),
        "i16_tile_unsorted": dict(move_kernel_ms=13.0, sort_tile=0),
        "i16_tile": dict(move_kernel_ms=13.0, sort_tile=0),
        "i16_tile_euler": dict(move_kernel_ms=8.0, sort_tile=0),
        "bf16_unsorted": dict(move_kernel_ms=11.0, sort_tile=0), "bf16": dict(move_kernel_ms=9.0, sort_tile=1),
        "bf16_lds": dict(move_kernel_ms=16.0, sort_tile=1)}
    line = {"steps": steps, "warmup": warmup, "value": 1.0, "variants": variants,
            "config": {"cell_tile_sort": 1, "sort_autotuned": True},
            "roofline": {"avg_launch_ms": 24.0, "frac": 0.5, "counters": {"key": "f64@t1"}}}
    (tmp_path / "bench.json").write_text("progress\n" + json.dumps(line) + "\n")
    md = tmp_path / "out.md"
    C.main(["bench-trace", "--bench", str(tmp_path / "bench.json"), "--moves", str(tmp_path / "moves.csv"),
            "--md", str(md)])
    rows = {r.split("|")[1].strip(): r for r in md.read_text().splitlines() if r.startswith("| ")}
    for key, want in ms.items():
        r = rows[key].split("|")
        assert float(r[4]) == want and float(r[5]) == want, (key, r)      # bench ms, trace ms
    assert "tune" not in rows


def test_assign_rejects_unaccounted_dispatches():
    runs = [dict(key="f64@t0", variant="f64", warmup_launches=1, timed_launches=2)]
    groups = {"k": [{"ns": 1}] * 4}
    try:
        C.assign(runs, groups)
    except SystemExit as e:
        assert "account for 3" in str(e)
    else:
        raise AssertionError("a dispatch count that the runs do not explain must be refused")


def test_short_names():
    assert C.short_name(NAMES["i16_tile"]) == "move_fast_kernel<short, TiledI16, false, false>"
    mangled = "_ZN2wc12_GLOBAL__N_116move_fast_kernelIDF16_NS0_11TiledPackedIDF16_EELb0ELb0EEEvNS_10MoveParamsE"
    assert C.short_name(mangled) == "move_fast_kernel<f16, TiledPacked<f16>>"

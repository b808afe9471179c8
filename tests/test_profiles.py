"""The committed move-kernel counters (profiles/move_counters.json) are what
bench.py's roofline reads: every run key the bench can name must be there
with one value per interval of the 24-hour window, the copy calibration must
be a plausible bytes-per-unit, and the issue rate they imply at the launch
times recorded beside them must stay below the chip's VALU issue peak."""
import json
import os

import numpy as np
import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TILES = (0, 1, 2, 4, 8)          # wc_autotune_sort's candidates


@pytest.fixture(scope="module")
def counters():
    with open(os.path.join(ROOT, "profiles", "move_counters.json")) as f:
        return json.load(f)


def keys():
    out = []
    for v in ("f64",) + bench.STORES:
        for t in TILES:
            out.append(f"{v}@t{t}")
            if v in bench.EULER_VARIANTS:
                out.append(f"{v}_euler@t{t}")
    return out


def test_every_bench_key_present(counters):
    k = counters["kernels"]
    for key in keys():
        assert key in k, key
        e = k[key]
        for f in ("valu_per_interval", "mem_bytes_per_interval", "fetch_bytes_per_interval",
                  "write_bytes_per_interval"):
            assert len(e[f]) == 24, (key, f)
        assert all(x > 0 for x in e["valu_per_interval"])


def test_calibration_plausible(counters):
    cal = counters["calibration"]
    for side in ("fetch", "write"):
        c = cal[side]
        assert c["copy_bytes"] == 1 << 30
        assert 256 <= c["bytes_per_unit"] <= 8192, c        # the counters read in KiB-scale units


def test_issue_fraction_below_peak(counters):
    for key, e in counters["kernels"].items():
        valu = np.mean(e["valu_per_interval"])
        frac = valu / (e["bench_launch_ms"] / 1e3) / bench.VALU_ISSUE_PEAK
        assert 0.2 < frac <= 1.0, (key, frac)
        # unsorted launches move the state coalesced: about 96 B per particle
        if key.endswith("@t0"):
            mem = np.mean(e["mem_bytes_per_interval"])
            assert 0.8 * 96 * 4e6 < mem < 1.2 * 96 * 4e6, (key, mem)


def test_roofline_block_reads_counters(counters):
    e = counters["kernels"]["f64@t1"]
    r = dict(move_ms=24 * e["bench_launch_ms"], move_launches=24, steps=4.0e6 * 120 * 24)
    b = bench.roofline_block("f64@t1", r, 4_000_000, 4, 8, True, counters["kernels"])
    assert b["counters"]["key"] == "f64@t1"
    assert b["frac"] == pytest.approx(np.mean(e["valu_per_interval"]) / (e["bench_launch_ms"] / 1e3)
                                      / bench.VALU_ISSUE_PEAK)
    assert b["hbm"]["frac"] < 1.0
    # a run shorter than the window averages its own intervals only
    r20 = dict(move_ms=20 * e["bench_launch_ms"], move_launches=20, steps=4.0e6 * 120 * 20)
    b20 = bench.roofline_block("f64@t1", r20, 4_000_000, 4, 8, True, counters["kernels"])
    assert b20["counters"]["valu_per_dispatch"] == pytest.approx(np.mean(e["valu_per_interval"][:20]))
    # an unknown key leaves the roofline unmeasured rather than borrowing another run's counters
    assert bench.roofline_block("f64@t16", r, 4_000_000, 4, 8, True, counters["kernels"])["frac"] is None

"""GPU tests of the cell-tile order of the context path (wc_set_sort): the
moves and the gridding visit the particles tile by tile through a
permutation kept beside the population.  Every result is per particle (the
draws are keyed by id, counts are integer sums), so the sorted runs must give
the unsorted runs' population bit for bit, in the same (reference) order,
through compaction, appends and both move paths; and the autotuner must leave
the population and the step count as it found them."""
import numpy as np
import pytest

import wacommplusplus_amd as wc
from runner import load_grid
from wacommplusplus_amd.synthetic import full_window, release

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full():
    return full_window(T=7)


def run(ctx, W, cfg, p, n, path, store, tile, period, extra=None):
    """n intervals with removal of the dead and gridding after each; `extra`
    particles are appended after the second interval."""
    load_grid(ctx, W, W.T, store)
    for t in range(W.T):
        ctx.upload_record(t, *W.record(t))
    ctx.set_sort(tile, period)
    ctx.set_particles(p)
    frames = []
    for s in range(n):
        ctx.move_interval(cfg, float(W.times[s]), float(W.times[s + 1]), s, s + 1, path=path)
        ctx.remove_dead()
        frames.append(ctx.grid())
        if s == 1 and extra is not None:
            ctx.append_particles(extra)
    out = ctx.get_particles()
    ctx.set_sort(0)
    return out, frames


@pytest.mark.parametrize("path,store", [("parity", "f32"), ("fast", "bf16_tile"), ("fast", "bf16"),
                                        ("fast", "i16_tile")])
def test_sorted_moves_equal_unsorted(full, path, store):
    p = release(full, 4_000_000, 50, spread=2.0)[::20].copy()
    extra = release(full, 4_000_000, 50, spread=2.0)[7::400].copy()
    extra["id"] += 10_000_000
    extra["time"] = float(full.times[2])
    cfg = wc.default_config(random=True, sigma=wc.sigma_from_kh(1.0, 30.0), horizontalClosure=wc.KILL)
    with wc.Context(0) as ctx:
        ref, ref_frames = run(ctx, full, cfg, p, 6, path, store, 0, 1, extra)
        assert len(ref) < len(p) + len(extra)                      # the kill closure removed some
        for tile, period in ((1, 1), (2, 3), (8, 1)):
            got, frames = run(ctx, full, cfg, p, 6, path, store, tile, period, extra)
            for f in ("id", "k", "j", "i", "health", "age", "time"):
                assert np.array_equal(got[f], ref[f], equal_nan=f != "id"), (tile, period, f)
            for a, b in zip(frames, ref_frames):
                assert np.array_equal(a, b), (tile, period)


def test_sorted_run_window_equals_unsorted():
    W = full_window(T=5, eta=256, xi=192)
    cfg = wc.default_config(random=True, sigma=wc.sigma_from_kh(1.0, 30.0))
    src = [wc.Source(k=-0.3, j=100.5, i=80.5, mode=wc.UNIFORM_RATE, value=20000.0, start=0.0, end=-1.0),
           wc.Source(k=-1.2, j=40.2, i=150.7, mode=wc.SINGLE_PULSE, value=30000.0, start=0.0, end=-1.0)]
    out = {}
    with wc.Context(0) as ctx:
        for tile in (0, 2):
            load_grid(ctx, W, 2)
            ctx.set_sort(tile, 1)
            ctx.set_particles(wc.make_particles(0))
            conc, stats = ctx.run_window(cfg, W.times, *W.stacked(), sources=src, path="parity")
            out[tile] = (conc, ctx.get_particles())
    assert np.array_equal(out[0][0], out[2][0])
    for f in ("id", "k", "j", "i", "health", "age", "time"):
        assert np.array_equal(out[0][1][f], out[2][1][f], equal_nan=f != "id"), f


def test_autotune_sort_restores_population(full):
    p = release(full, 4_000_000, 50, spread=2.0)[::10].copy()
    cfg = wc.default_config(random=True, sigma=wc.sigma_from_kh(1.0, 30.0))
    with wc.Context(0) as ctx:
        load_grid(ctx, full, full.T, "bf16_tile")
        for t in range(full.T):
            ctx.upload_record(t, *full.record(t))
        ctx.set_particles(p)
        ctx.move_interval(cfg, float(full.times[0]), float(full.times[1]), 0, 1, path="fast")
        before = ctx.get_particles()
        steps = ctx.step_count()
        tile, ms = ctx.autotune_sort(cfg, float(full.times[1]), float(full.times[2]), 1, 2, path="fast")
        assert tile in (0, 1, 2, 4, 8) and ms > 0
        assert ctx.get_sort() == (tile, 1)
        after = ctx.get_particles()
        for f in ("id", "k", "j", "i", "health", "age", "time"):
            assert np.array_equal(after[f], before[f], equal_nan=f != "id"), f
        assert ctx.step_count() == steps
        print(f"autotuned sort tile {tile} ({ms:.2f} ms on one interval of {len(p)} particles)", flush=True)

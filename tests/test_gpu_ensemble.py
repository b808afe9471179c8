"""GPU tests of the batched ensemble (wc_ens_*): many release scenarios,
each with its own source, horizontal diffusion scale and release offset,
moved as one population on the mixed-precision path from one captured step.

The ensemble has no counterpart in the reference, so it is pinned to the
mixed-precision context path (itself held to the FP64 path by
tests/test_gpu_fast.py):
  - one scenario with the config's sigma gives the context path's positions
    bit for bit (same operator, same Philox draws keyed on id, substep and
    interval; health differs only by its FP32 storage);
  - per-scenario sigma: the displacement variance of each scenario after one
    interval under zero current is 120 sigma_s^2 (docs/model.md:204-210);
  - graph replays equal direct launches after every interval, with other
    work and synchronisation between the replays;
  - the surface counts equal a brute-force count of the downloaded state;
  - particles whose release offset lies beyond the interval neither move
    nor count;
  - autotuning leaves the population unchanged and the results identical.
"""
import math

import numpy as np
import pytest

import wacommplusplus_amd as wc
from runner import run_device
from wacommplusplus_amd.synthetic import ForcingWindow, ensemble, full_window, rectilinear, release

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = wc.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def full():
    return full_window(T=5)


def load_window(ctx, W, store="bf16"):
    ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask, n_slots=W.T, store=store)
    for t in range(W.T):
        ctx.upload_record(t, *W.record(t))


def run_ens(ctx, W, cfg, n_intervals, use_graph=True, surface=True, first=0):
    times = [float(W.times[first + q]) for q in range(n_intervals + 1)]
    slots = [first + q for q in range(n_intervals + 1)]
    return ctx.ens_run(cfg, times, slots, use_graph=use_graph, surface=surface)


@pytest.mark.parametrize("random", [False, True])
def test_one_scenario_equals_fast_path(ctx, full, random):
    p = release(full, 4_000_000, 50, spread=0.25)[::200].copy()
    cfg = wc.default_config(random=random, sigma=wc.sigma_from_kh(1.0, 30.0))
    ref = run_device(ctx, full, cfg, p, 4, path="fast", store="bf16")
    load_window(ctx, full)
    ctx.ens_set(p, np.zeros(len(p), np.uint8), [cfg.sigma])
    st = run_ens(ctx, full, cfg, 4)
    got, sc = ctx.ens_get()
    assert np.array_equal(got["id"], ref["id"]) and np.all(sc == 0)
    for f in ("k", "j", "i", "age"):
        assert np.array_equal(got[f], ref[f], equal_nan=True), f
    assert np.array_equal(got["health"] > 0, ref["health"] > 0)
    alive = ref["health"] > 0
    assert np.max(np.abs(got["health"][alive] - ref["health"][alive])) < 1e-7
    assert st.particle_steps > 0.9 * len(p) * 4 * 120


def still_window(T=2, eta=64, xi=64, s_rho=4):
    lon, lat = rectilinear(eta, xi, 150.0)
    s_w = s_rho + 1

    def rec(t):
        return (np.zeros((eta, xi), np.float32), np.zeros((s_rho, eta, xi), np.float32),
                np.zeros((s_rho, eta, xi), np.float32), np.zeros((s_w, eta, xi), np.float32),
                np.zeros((s_w, eta, xi), np.float32))
    return ForcingWindow(s_rho, eta, xi, lon, lat, np.full((eta, xi), 40.0), np.ones((eta, xi)),
                         np.full(s_w - 1, 1.0 / (s_w - 1)), np.arange(T) * 3600.0, rec)


def test_per_scenario_sigma(ctx):
    W = still_window()
    load_window(ctx, W, store="f32")
    n = 100_000
    kh = np.array([0.1, 1.0, 10.0])
    sigma = np.sqrt(2 * kh * 30.0)
    p = wc.make_particles(3 * n, k=-1.5, j=32.25, i=32.75)
    p["id"] = np.arange(3 * n)
    scen = np.repeat(np.arange(3, dtype=np.uint8), n)
    ctx.ens_set(p, scen, sigma)
    cfg = wc.default_config(random=True, sigma=123.0)       # unused: sigma comes from the scenario
    run_ens(ctx, W, cfg, 1)
    out, sc = ctx.ens_get()
    dx_tab, dy_tab = wc.cell_metric_table(W.lonRad, W.latRad)
    kLL = -W.s_w + 2
    se = math.sqrt(2.0 / n)
    for s in range(3):
        m = sc == s
        dx = (out["i"][m] - 32.75) * dx_tab[32, 32]
        dy = (out["j"][m] - 32.25) * dy_tab[32, 32]
        expected = 120 * (sigma[s] * (1 - (-1.5) / kLL)) ** 2
        assert abs(dx.var() / expected - 1) < 5 * se, (s, dx.var() / expected)
        assert abs(dy.var() / expected - 1) < 5 * se, (s, dy.var() / expected)


def brute_surface(out, sc, n_scen, eta, xi, t_end):
    counts = np.zeros((n_scen, eta, xi), np.uint32)
    # rounded level 0 with half away from zero: k > -0.5 (k <= 0 after the closures)
    ok = (out["health"] > 0) & (out["k"] > -0.5) & (out["time"] <= t_end) & np.isfinite(out["j"])
    jr = np.floor(out["j"][ok] + 0.5).astype(np.int64)
    ir = np.floor(out["i"][ok] + 0.5).astype(np.int64)
    inside = (jr < eta) & (ir < xi)
    np.add.at(counts, (sc[ok][inside].astype(np.int64), jr[inside], ir[inside]), 1)
    return counts


def test_graph_replays_equal_direct_launches_and_surface_counts(ctx, full):
    p, scen, sigma, counts = ensemble(full, n_scenarios=32, total=400_000, median=10_000, lo=2_000, hi=60_000,
                                      max_offset_h=2.0)
    cfg = wc.default_config(random=True, sigma=1.0)
    load_window(ctx, full)
    results = {}
    for use_graph in (False, True):
        ctx.ens_set(p, scen, sigma)
        initial, _ = ctx.ens_get()
        results[use_graph] = []
        for q in range(4):                      # one replay per call, host work between
            run_ens(ctx, full, cfg, 1, use_graph=use_graph, first=q)
            out, sc = ctx.ens_get()
            surf = ctx.ens_surface()
            results[use_graph].append((out, sc, surf))
    for q in range(4):
        a, b = results[False][q], results[True][q]
        for f in ("k", "j", "i", "health", "age"):
            assert np.array_equal(a[0][f], b[0][f], equal_nan=True), (q, f)
        assert np.array_equal(a[2], b[2]), q
        out, sc, surf = b
        t_end = float(full.times[q + 1])
        assert np.array_equal(surf, brute_surface(out, sc, len(sigma), full.eta, full.xi, t_end)), q
    # particles released after the first interval did not move in it
    out0 = results[True][0][0]
    late = p["time"] > float(full.times[1])
    assert late.any()
    for f in ("k", "j", "i", "age", "health"):
        assert np.array_equal(out0[f][late], initial[f][late]), f
    assert np.all(out0["age"][late] == 0)


def test_multi_interval_graph_run_and_autotune(ctx, full):
    p, scen, sigma, counts = ensemble(full, n_scenarios=16, total=200_000, median=10_000, lo=2_000, hi=40_000,
                                      max_offset_h=1.0)
    cfg = wc.default_config(random=True, sigma=1.0)
    load_window(ctx, full)
    ctx.ens_set(p, scen, sigma)
    ctx.ens_set_launch(256, 1, 0)
    run_ens(ctx, full, cfg, 4, use_graph=False)
    ref, _ = ctx.ens_get()
    ref_surf = ctx.ens_surface()
    ctx.ens_set(p, scen, sigma)
    initial, _ = ctx.ens_get()
    best = ctx.ens_autotune(cfg, float(full.times[0]), float(full.times[1]), 0, 1)
    shape = ctx.ens_get_launch()
    print(f"autotuned launch {shape}: {best:.3f} ms", flush=True)
    assert best > 0 and shape[0] in (64, 128, 256) and shape[1] in (1, 2, 4) and shape[2] in (0, 3, 4)
    assert ctx.ens_get_sort()[0] in (0, 2, 4, 8) and ctx.ens_get_lds()[0] in (0, 10, 30)
    after, _ = ctx.ens_get()                     # the population is restored by the autotuner
    for f in ("k", "j", "i", "health", "age"):
        assert np.array_equal(after[f], initial[f], equal_nan=True), f
    st = run_ens(ctx, full, cfg, 4, use_graph=True)
    got, _ = ctx.ens_get()
    for f in ("k", "j", "i", "health", "age"):
        assert np.array_equal(got[f], ref[f], equal_nan=True), f
    assert np.array_equal(ctx.ens_surface(), ref_surf)
    assert st.intervals == 4 and st.device_ms > 0


def test_permuted_population_same_per_id(ctx, full):
    p, scen, sigma, counts = ensemble(full, n_scenarios=8, total=100_000, median=10_000, lo=2_000, hi=30_000)
    cfg = wc.default_config(random=True, sigma=1.0)
    load_window(ctx, full)
    ctx.ens_set(p, scen, sigma)
    run_ens(ctx, full, cfg, 2)
    a, sa = ctx.ens_get()
    perm = np.random.default_rng(5).permutation(len(p))
    ctx.ens_set(p[perm], scen[perm], sigma)
    run_ens(ctx, full, cfg, 2)
    b, sb = ctx.ens_get()
    order = np.argsort(b["id"])
    for f in ("k", "j", "i", "health", "age"):
        assert np.array_equal(b[f][order], a[f], equal_nan=True), f
    assert np.array_equal(sb[order], sa)


@pytest.mark.parametrize("tile,period", [(2, 1), (4, 1), (8, 3)])
def test_cell_tile_sort_changes_no_result(ctx, full, tile, period):
    """Sorting the population by cell tile (in the captured step) gives every
    particle the same state and the same surface counts as no sorting."""
    p, scen, sigma, counts = ensemble(full, n_scenarios=16, total=200_000, median=10_000, lo=2_000, hi=40_000)
    cfg = wc.default_config(random=True, sigma=1.0)
    load_window(ctx, full)
    res = []
    for t in (0, tile):
        ctx.ens_set(p, scen, sigma)
        ctx.ens_set_sort(t, period)
        run_ens(ctx, full, cfg, 4, use_graph=True)
        res.append((ctx.ens_get(), ctx.ens_surface()))
    (a, sa), ga = res[0]
    (b, sb), gb = res[1]
    for f in ("id", "k", "j", "i", "health", "age", "time"):
        assert np.array_equal(a[f], b[f], equal_nan=True), f
    assert np.array_equal(sa, sb) and np.array_equal(ga, gb)
    ctx.ens_set_sort(0, 1)


@pytest.mark.parametrize("chunk,halo,tile,occ", [(1, 2, 0, 0), (7, 2, 4, 3), (30, 0, 0, 3), (30, 2, 4, 0),
                                                 (120, 16, 2, 3)])
def test_lds_tiled_move_changes_no_result(ctx, full, chunk, halo, tile, occ):
    """The LDS-tiled move (forcing staged per workgroup, restaged every `chunk`
    substeps; halo 0 sends many stencils to global memory) gives every
    particle the plain move's state and surface counts, bit for bit."""
    p, scen, sigma, counts = ensemble(full, n_scenarios=16, total=200_000, median=10_000, lo=2_000, hi=40_000,
                                      max_offset_h=1.0)
    cfg = wc.default_config(random=True, sigma=1.0)
    load_window(ctx, full)
    res = []
    for lds in (0, chunk):
        ctx.ens_set(p, scen, sigma)
        ctx.ens_set_lds(0, 2)
        ctx.ens_set_launch(256, 1, occ)
        ctx.ens_set_sort(tile, 1)
        ctx.ens_set_lds(lds, halo)
        assert ctx.ens_get_lds() == (lds, halo)
        st = run_ens(ctx, full, cfg, 3, use_graph=True)
        res.append((ctx.ens_get(), ctx.ens_surface(), st.particle_steps))
    (a, sa), ga, na = res[0]
    (b, sb), gb, nb = res[1]
    for f in ("id", "k", "j", "i", "health", "age", "time"):
        assert np.array_equal(a[f], b[f], equal_nan=True), f
    assert np.array_equal(sa, sb) and np.array_equal(ga, gb) and na == nb
    ctx.ens_set_lds(0, 2)
    ctx.ens_set_sort(0, 1)


def test_rejects_bad_input(ctx, full):
    load_window(ctx, full)
    p = wc.make_particles(10, j=10.5, i=10.5)
    with pytest.raises(wc.WcError):
        ctx.ens_set(p, np.full(10, 3, np.uint8), [1.0, 2.0])          # scenario >= n_scenarios
    with pytest.raises(wc.WcError):
        ctx.ens_set_launch(100, 1, 0)
    with pytest.raises(wc.WcError):
        ctx.ens_set_lds(-1, 2)
    ctx.ens_set_launch(256, 1, 0)
    ctx.ens_set_lds(10, 2)
    with pytest.raises(wc.WcError):
        ctx.ens_set_launch(256, 1, 4)                                   # the LDS move needs hint 0 or 3
    ctx.ens_set_lds(0, 2)
    cfg = wc.default_config(driftModel=1)
    ctx.ens_set(p, np.zeros(10, np.uint8), [1.0])
    with pytest.raises(wc.WcError) as e:
        run_ens(ctx, full, cfg, 1)
    assert e.value.status == wc.WC_ERR_UNSUPPORTED

"""GPU tests of the mixed-precision path (WC_PATH_FAST): integer cell + FP32
offset state, FP32 interpolation of FP32/FP16/BF16 forcing (plain, or as
residuals from 4 x 4-cell tile means), Philox4x32-10 diffusion draws.  It is
not draw-for-draw identical to the reference, so it is held to stated
precision and statistical bars:

  - diffusion off, against the FP64 parity path on the same forcing, after
    24 h on a 40,000-particle sample of the full-size release (each particle
    offset by N(0, 0.25^2) cells so every trajectory is distinct).  The bar of
    the north star, <= 1e-3 cells, is asserted at the median and as RMS over
    the open-water particles: those whose trajectory, on either path, met
    neither a closure (a shore or open-boundary closure acted, or the
    midpoint fell back to Euler at land: WC_EVENT_SHORE | DOMAIN | FALLBACK)
    nor a stencil with a land node (WC_EVENT_COAST), where the reference's
    bilinear operator jumps to the land zeros and the trajectory is decided
    by which side of a node line a rounding puts it.  Without diffusion about
    80 % of this release meets one of them within 24 h (reported apart, as
    are the closure-free particles).  The RMS is a tail statistic:
    trajectories in this 64-mode flow separate exponentially, so it is
    asserted where it is met and marked xfail with the measured value where
    it is not.
  - diffusion on: the concentration field of the fast path differs from the
    FP64 path's by no more than two FP64 runs with different seeds differ
    (relative L2 of the 24 h count grids).
  - diffusion on, zero current: per-axis displacement variance after one
    interval = 120 sigma_k^2 (docs/model.md:204-210) within 5 standard errors,
    mean within 5 standard errors of zero, x/y draws uncorrelated.
"""
import math

import numpy as np
import pytest

import wacommplusplus_amd as wc
from runner import run_device
from wacommplusplus_amd.synthetic import ForcingWindow, full_window, plain_window, rectilinear, release

pytestmark = pytest.mark.gpu

PINNED = wc.EVENT["shore"] | wc.EVENT["domain"] | wc.EVENT["fallback"]
COASTAL = PINNED | wc.EVENT["coast"]


@pytest.fixture(scope="module")
def ctx():
    c = wc.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def full():
    return full_window()


def run_tracked(ctx, W, cfg, p, n, **kw):
    ctx.set_event_tracking(True)
    try:
        out = run_device(ctx, W, cfg, p, n, **kw)
        return out, ctx.get_events()
    finally:
        ctx.set_event_tracking(False)


@pytest.fixture(scope="module")
def reference_24h(ctx, full):
    # distinct trajectories: each particle offset by N(0, 0.25^2) cells
    p = release(full, 4_000_000, 50, spread=0.25)[::100].copy()
    cfg = wc.default_config(random=False)
    ref, ev = run_tracked(ctx, full, cfg, p, 24, path="parity")
    return p, cfg, ref, ev


def deviation(a, b, ev_a=None, ev_b=None):
    """Per-particle max(|dk|, |dj|, |di|) over the particles alive in both;
    with events, also over the ones no closure touched on either path."""
    both = (a["health"] > 0) & (b["health"] > 0)
    d = np.max(np.stack([np.abs(a[f] - b[f]) for f in ("k", "j", "i")]), axis=0)

    def stats(sel):
        x = d[sel]
        q = np.quantile(x, [0.5, 0.9, 0.99])
        return dict(n=int(sel.sum()), p50=float(q[0]), p90=float(q[1]), p99=float(q[2]), max=float(x.max()),
                    rms=float(np.sqrt(np.mean(x * x))))
    r = stats(both)
    r["both_alive"] = float(both.mean())
    if ev_a is not None:
        free = both & (((ev_a | ev_b) & PINNED) == 0)
        r["free"] = stats(free)
        r["free_fraction"] = float(free.sum() / both.sum())
        open_water = both & (((ev_a | ev_b) & COASTAL) == 0)
        r["open"] = stats(open_water)
        r["open_fraction"] = float(open_water.sum() / both.sum())
    return r


# measured on the 4M release (median / RMS over the open-water particles):
# f32 4.8e-7 / 1.6e-4, f16_tile 6.4e-5 / 1.8e-3, bf16_tile 5.6e-4 / 5.6e-3,
# f16 5.8e-4 / 2.5e-3, bf16 3.2e-3 / 2.7e-2
BAR = 1e-3
CASES = [
    ("f32", True, True),
    ("i16_tile", True, True),
    ("f16_tile", True, False),
    ("bf16_tile", True, False),
    ("f16", True, False),
    ("bf16", False, False),
]


@pytest.mark.parametrize("store,median_met,rms_met", CASES)
def test_fast_path_deviation_vs_fp64(ctx, full, reference_24h, store, median_met, rms_met, request):
    p, cfg, ref, ev_ref = reference_24h
    dev, ev = run_tracked(ctx, full, cfg, p, 24, path="fast", store=store)
    r = deviation(dev, ref, ev, ev_ref)
    print(f"fast/{store} vs FP64, 24 h, diffusion off: {r}", flush=True)
    assert r["both_alive"] > 0.75                      # closures decide the same deaths
    assert 0.1 < r["open_fraction"] < r["free_fraction"] < 0.5
    ow = r["open"]
    if not median_met:
        request.applymarker(pytest.mark.xfail(reason=f"{store}: median {ow['p50']:.2e} cells > 1e-3 (8-bit "
                                                     "significand without tile means)", strict=False))
    elif not rms_met:
        request.applymarker(pytest.mark.xfail(reason=f"{store}: RMS {ow['rms']:.2e} cells > 1e-3 over the "
                                                     f"open-water particles (median {ow['p50']:.2e})",
                                              strict=False))
    assert ow["p50"] <= BAR
    assert ow["rms"] <= BAR


def test_tile_means_keep_zeros_exact(ctx):
    """Zero current and w everywhere: the tile-mean storage types decode the
    zeros exactly, so the particles end where FP32 storage leaves them (their
    FP32 offsets, unmoved)."""
    W = still_window()
    p = wc.make_particles(1000, k=-1.5, j=32.25, i=32.75)
    p["j"] += np.linspace(-5, 5, 1000)
    cfg = wc.default_config(random=False)
    ref = run_device(ctx, W, cfg, p, 1, path="fast", store="f32")
    assert np.max(np.abs(ref["j"] - p["j"])) < 1e-6
    for store in ("bf16_tile", "f16_tile", "i16_tile"):
        out = run_device(ctx, W, cfg, p, 1, path="fast", store=store)
        for f in ("k", "j", "i"):
            assert np.array_equal(out[f], ref[f]), (store, f)


def test_tile_means_resolve_uniform_current(ctx):
    """Plain configuration (uniform u = 0.2, v = 0.1, 200 m cells), 12 h:
    j(t) = j0 + v t / dy exactly in FP64.  A tile mean holds the value's top
    bits and the residual the rest, so BF16 tile storage lands within 2e-4
    cells of the closed form (measured 2.1e-5); plain BF16 rounds 0.1 to
    0.10009765625 and misses by 21.6 cells x 9.8e-4 = 0.021 cells."""
    W = plain_window(T=13)
    p = release(W, 500, 1, depth_m=1.0)
    p["j"], p["i"] = 59.5, 49.5
    cfg = wc.default_config(random=False)
    _, dy = wc.cell_metric_table(W.lonRad, W.latRad)
    j_exact = 59.5 + 0.1 * 12 * 3600.0 / dy[59, 49]
    tile = run_device(ctx, W, cfg, p, 12, path="fast", store="bf16_tile")
    plain = run_device(ctx, W, cfg, p, 12, path="fast", store="bf16")
    scaled = run_device(ctx, W, cfg, p, 12, path="fast", store="i16_tile")
    e_tile = float(np.max(np.abs(tile["j"] - j_exact)))
    e_plain = float(np.max(np.abs(plain["j"] - j_exact)))
    e_scaled = float(np.max(np.abs(scaled["j"] - j_exact)))
    print(f"uniform current, 12 h: |j - exact| bf16_tile {e_tile:.2e}, bf16 {e_plain:.2e}, i16_tile {e_scaled:.2e}",
          flush=True)
    assert e_tile < 2e-4 and e_plain > 0.01
    # a uniform tile keeps its mean to FP32 precision (residuals 0)
    assert e_scaled < 2e-4


def counts_24h(ctx, full, p, cfg, path, store):
    run_device(ctx, full, cfg, p, 24, path=path, store=store)
    return ctx.grid().astype(np.float64)


def test_fast_path_concentration_within_seed_noise(ctx, full):
    """Relative L2 difference of the 24 h count grid (diffusion on, 400k
    particles): fast path vs FP64 path, against FP64 seed A vs FP64 seed B."""
    p = release(full, 4_000_000, 50, spread=0.25)[::10].copy()
    cfg_a = wc.default_config(random=True, sigma=wc.sigma_from_kh(1.0, 30.0), randomSeed=5489)
    cfg_b = cfg_a.copy(randomSeed=1234)
    ref_a = counts_24h(ctx, full, p, cfg_a, "parity", "f32")
    ref_b = counts_24h(ctx, full, p, cfg_b, "parity", "f32")
    norm = np.linalg.norm(ref_a)
    noise = np.linalg.norm(ref_a - ref_b) / norm
    for store in ("bf16", "f16", "f32"):
        got = counts_24h(ctx, full, p, cfg_a, "fast", store)
        err = np.linalg.norm(got - ref_a) / norm
        print(f"conc L2 fast/{store} vs FP64 = {err:.4f}; FP64 seed-to-seed = {noise:.4f}", flush=True)
        assert err <= 1.2 * noise


def still_window(T=2, eta=64, xi=64, s_rho=4):
    """Zero current, AKt=0: only the horizontal random walk moves particles."""
    lon, lat = rectilinear(eta, xi, 150.0)
    s_w = s_rho + 1

    def rec(t):
        return (np.zeros((eta, xi), np.float32), np.zeros((s_rho, eta, xi), np.float32),
                np.zeros((s_rho, eta, xi), np.float32), np.zeros((s_w, eta, xi), np.float32),
                np.zeros((s_w, eta, xi), np.float32))
    return ForcingWindow(s_rho, eta, xi, lon, lat, np.full((eta, xi), 40.0), np.ones((eta, xi)),
                         np.full(s_w - 1, 1.0 / (s_w - 1)), np.arange(T) * 3600.0, rec)


@pytest.mark.parametrize("store", ["f32", "bf16"])
def test_fast_path_diffusion_statistics(ctx, store):
    W = still_window()
    n = 200_000
    p = wc.make_particles(n, k=-1.5, j=32.25, i=32.75)
    sigma = wc.sigma_from_kh(1.0, 30.0)
    cfg = wc.default_config(random=True, sigma=sigma, integrator=wc.MIDPOINT)
    out = run_device(ctx, W, cfg, p, 1, path="fast", store=store)
    dx_tab, dy_tab = wc.cell_metric_table(W.lonRad, W.latRad)
    dx_m = (out["i"] - 32.75) * dx_tab[32, 32]
    dy_m = (out["j"] - 32.25) * dy_tab[32, 32]
    kLL = -W.s_w + 2
    sigma_k = sigma * (1 - (-1.5) / kLL)
    expected = 120 * sigma_k ** 2
    se = math.sqrt(2.0 / n)
    assert abs(dx_m.var() / expected - 1) < 5 * se
    assert abs(dy_m.var() / expected - 1) < 5 * se
    assert abs(dx_m.mean()) < 5 * math.sqrt(expected / n)
    assert abs(dy_m.mean()) < 5 * math.sqrt(expected / n)
    assert abs(np.corrcoef(dx_m, dy_m)[0, 1]) < 5 / math.sqrt(n)
    assert np.all(out["k"] == -1.5)                           # AKt = 0: no vertical random leap


def test_fast_path_uniform_current_matches_parity(ctx):
    """Plain configuration (uniform u=0.2, v=0.1 exactly representable as the
    same float in every storage): FP32 fast path within 1e-4 cells of the
    FP64 path after 12 h."""
    W = plain_window(T=13)
    p = release(W, 2000, 1, depth_m=1.0)
    p["j"], p["i"] = 59.5, 49.5
    cfg = wc.default_config(random=False)
    ref = run_device(ctx, W, cfg, p, 12, path="parity")
    dev = run_device(ctx, W, cfg, p, 12, path="fast", store="f32")
    r = deviation(dev, ref)
    assert r["both_alive"] == 1.0 and r["max"] < 1e-4, r


def test_fast_path_backward_matches_parity(ctx):
    """Backward tracking (direction -1 from the last record) on the plain
    configuration: the FP32 fast path within 1e-4 cells of the FP64 path after
    12 h, and a forward run followed by the backward run returns every
    particle to its release position: within 1e-6 cells on the FP64 path
    (the oracle's own return error here is 6.5e-8: the row-dependent cell
    width makes the midpoint return inexact at second order) and within
    1e-5 cells on the FP32 path (measured 2.4e-7)."""
    W = plain_window(T=13)
    p = release(W, 2000, 1, depth_m=1.0)
    p["j"], p["i"] = 59.5, 49.5
    p["time"] = float(W.times[-1])
    back = wc.default_config(random=False, trackingDirection=-1)
    ref = run_device(ctx, W, back, p, 12, path="parity")
    dev = run_device(ctx, W, back, p, 12, path="fast", store="f32")
    r = deviation(dev, ref)
    assert r["both_alive"] == 1.0 and r["max"] < 1e-4, r
    assert np.all(ref["j"] < 59.5) and np.all(ref["i"] < 49.5)        # moved against the current
    q = p.copy()
    q["time"] = float(W.times[0])
    for path, store, tol in (("parity", "f32", 1e-6), ("fast", "f32", 1e-5)):
        fwd = run_device(ctx, W, wc.default_config(random=False), q, 12, path=path, store=store)
        fwd["time"] = float(W.times[-1])
        ret = run_device(ctx, W, back, fwd, 12, path=path, store=store)
        for f in ("j", "i"):
            err = float(np.max(np.abs(ret[f] - p[f])))
            print(f"{path}: return error in {f} {err:.2e} cells", flush=True)
            assert err < tol, (path, f, err)


def test_fast_path_full_size_invariants(ctx, full):
    p = release(full, 4_000_000, 50)
    cfg = wc.default_config(random=True, sigma=wc.sigma_from_kh(1.0, 30.0))
    a, frames = run_device(ctx, full, cfg, p, 2, path="fast", store="bf16", grid_frames=True)
    alive = a[a["health"] > 0]
    assert len(alive) > 0.99 * len(p)
    kr, jr, ir = np.round(alive["k"]), np.round(alive["j"]), np.round(alive["i"])
    inside = (jr >= 0) & (jr < full.eta) & (ir >= 0) & (ir < full.xi) & (kr >= -full.s_rho + 1) & (kr <= 0)
    assert int(frames[-1].sum()) == int(inside.sum())
    # keyed draws: a permuted population gives identical per-id results
    perm = np.random.default_rng(3).permutation(len(p))
    b = run_device(ctx, full, cfg, p[perm], 2, path="fast", store="bf16")
    b = b[np.argsort(b["id"])]
    for f in ("id", "k", "j", "i", "health", "age", "time"):
        assert np.array_equal(b[f], a[f], equal_nan=f != "id"), f

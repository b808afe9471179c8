"""GPU test of wc_grid_noalloc, the allocation-free gridding entry point:
its float frame equals wc_grid's (the context entry point whose count frame
the library allocates) bit for bit, with and without the land mask, on two
calls in a row -- the full-size 1024 x 768 x 11 frame with 1M particles, then
a smaller 96 x 64 x 5 frame with fewer particles, which must not see the
first call's counts left in the shared buffer -- and frames above the stated
maximum are refused."""
import numpy as np
import pytest

import wacommplusplus_amd as wc
from hip_rt import DeviceBuffer
from wacommplusplus_amd.synthetic import full_window, release

pytestmark = pytest.mark.gpu


def frame_case(eta, xi, s_rho, n, seed):
    """Positions spread over the frame with .5 ties, dead and out-of-domain
    particles, a land mask of ~30 %."""
    rng = np.random.default_rng(seed)
    p = wc.make_particles(n)
    p["k"] = rng.uniform(-s_rho + 0.4, 0.4, n)
    p["j"] = rng.uniform(-0.6, eta - 0.4, n)
    p["i"] = rng.uniform(-0.6, xi - 0.4, n)
    tie = rng.random(n) < 0.05
    p["j"][tie] = np.floor(p["j"][tie]) + 0.5
    p["k"][tie] = np.round(p["k"][tie]) - 0.5
    p["health"] = np.where(rng.random(n) < 0.1, -1.0, 1.0)
    mask = (rng.random((eta, xi)) > 0.3).astype(np.float64)
    lon = np.tile(np.linspace(0.2, 0.21, xi), (eta, 1))
    lat = np.tile(np.linspace(0.7, 0.71, eta)[:, None], (1, xi))
    return p, mask, lon, lat


def context_frame(p, mask, lon, lat, s_rho, mask_output):
    eta, xi = mask.shape
    with wc.Context(0) as ctx:
        ctx.set_grid(lon, lat, np.full((eta, xi), 40.0), np.full(s_rho, 1.0 / s_rho), s_rho, mask=mask,
                     n_slots=2)
        ctx.set_particles(p)
        return ctx.grid(counts=False, conc=True, mask_output=mask_output)


def noalloc_frame(p, mask, s_rho, mask_output):
    eta, xi = mask.shape
    L = wc.load_library()
    pd, md = DeviceBuffer(p), DeviceBuffer(mask)
    out = np.full((s_rho, eta, xi), -7.0, np.float32)        # overwritten entirely
    cd = DeviceBuffer(out)
    try:
        assert L.wc_grid_noalloc(pd.ptr, len(p), s_rho, eta, xi, md.ptr, int(mask_output), cd.ptr, None) == wc.WC_OK
        return cd.get(out)
    finally:
        pd.free()
        md.free()
        cd.free()


@pytest.mark.parametrize("mask_output", [False, True])
def test_noalloc_grid_matches_wc_grid(mask_output):
    for eta, xi, s_rho, n, seed in ((1024, 768, 11, 1_000_000, 1), (96, 64, 5, 20_000, 2)):
        p, mask, lon, lat = frame_case(eta, xi, s_rho, n, seed)
        want = context_frame(p, mask, lon, lat, s_rho, mask_output)
        got = noalloc_frame(p, mask, s_rho, mask_output)
        assert got.shape == want.shape
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (eta, xi, s_rho)
        live = (p["health"] > 0)
        assert 0 < float(want[want < 1e36].sum()) <= live.sum()


def test_noalloc_grid_full_release():
    """The full-size release after one interval of the FP64 move (clustered
    plumes, equal cells in neighbouring particles)."""
    W = full_window(T=2)
    p = release(W, 4_000_000, 50)
    cfg = wc.default_config(random=True, sigma=wc.sigma_from_kh(1.0, 30.0))
    with wc.Context(0) as ctx:
        ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask, n_slots=2)
        ctx.upload_record(0, *W.record(0))
        ctx.upload_record(1, *W.record(1))
        ctx.set_particles(p)
        ctx.move_interval(cfg, 0.0, 3600.0, 0, 1)
        moved = ctx.get_particles()
        want = ctx.grid(counts=False, conc=True, mask_output=True)
    got = noalloc_frame(moved, W.mask, W.s_rho, True)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_noalloc_grid_rejects_oversized_frames():
    L = wc.load_library()
    p = wc.make_particles(4)
    pd = DeviceBuffer(p)
    out = DeviceBuffer(np.zeros(16, np.float32))
    try:
        too_big = wc.WC_ERR_INVALID
        assert L.wc_grid_noalloc(pd.ptr, 4, 17, 1024, 1024, None, 0, out.ptr, None) == too_big   # 17.8M cells
        assert L.wc_grid_noalloc(pd.ptr, 4, 1, 4, 4, None, 1, out.ptr, None) == too_big        # mask required
        assert L.wc_grid_noalloc(None, 4, 1, 4, 4, None, 0, out.ptr, None) == too_big           # particles
    finally:
        pd.free()
        out.free()

"""Parity of the literal cudaMoveParticle replacement, wc_hip_move_particles
(cuda/kernel.h:7, kernel.cu:370-382): particle_data AoS in device memory, the
reference's device layout of the whole window (kernel.cu:63-69), one call per
interval -- against the oracle.  The cell metric is the haversine table of
Particle.cpp:544-579 computed on the host once per lon/lat grid (the grid's
content hash is taken on the device at every call), as the context path
does: recomputing it with the device sin/cos/atan2/pow differed from glibc
by ulps, which the reference's reflectCell (discontinuous where a candidate
crosses the edge of the old cell, NumericalHelpers.hpp:62-65) grew to 2e-4
cells in 3 h.

Bars of tests/CudaParticleParityTest.cu:141: 1e-12 cells deterministic
(every particle), 1e-9 stochastic.  With diffusion the device log/cos of the
draws may differ from glibc by an ulp, which can move a particle exactly at a
closure threshold to the other side; so, as in the context-path sweep, at
least 97 % per case and 99.9 % over all cases meet 1e-9."""
import numpy as np
import pytest

import wacommplusplus_amd as wc
from hip_rt import run_dropin
from runner import run_oracle
from test_gpu_parity import close, random_window
from wacommplusplus_amd.synthetic import full_window, release

pytestmark = pytest.mark.gpu


def test_dropin_random_windows_all_closures():
    det, rnd, bad = [], [], []
    for seed in range(6):
        rng = np.random.default_rng(100 + seed)
        W = random_window(seed)
        n = 400
        p = wc.make_particles(n)
        p["k"] = rng.uniform(-3.2, 0.2, n)
        p["j"] = rng.uniform(-0.5, W.eta - 0.5, n)
        p["i"] = rng.uniform(-0.5, W.xi - 0.5, n)
        p["time"] = rng.choice([0.0, 0.0, 900.0, 2000.0], n)
        p["id"] = rng.integers(0, 2**63, n, dtype=np.uint64)
        for direction in (1, -1):
            for integ in (1, 2):
                for closures in ((1, 1, 1), (2, 2, 2), (3, 3, 3), (1, 2, 3)):
                    for random in (False, True):
                        case = (seed, direction, integ, closures, random)
                        cfg = wc.default_config(integrator=integ, upperClosure=closures[0],
                                                lowerClosure=closures[1], horizontalClosure=closures[2],
                                                random=random, sigma=30.0, survprob=1e-4, tau0=20000.0,
                                                randomSeed=seed, trackingDirection=direction, backwardDiffusion=1)
                        ref = run_oracle(W, cfg, p, 3)
                        dev = run_dropin(W, cfg, p, 3)
                        tol = 1e-9 if random else 1e-12
                        ok = np.all(np.stack([close(dev[f], ref[f], tol) for f in ("k", "j", "i", "age")]), axis=0)
                        ok &= close(dev["health"], ref["health"], 1e-12)
                        (rnd if random else det).append(ok.mean())
                        if ok.mean() < (0.97 if random else 1.0):
                            bad.append((case, float(ok.mean())))
        print(f"window {seed}: done", flush=True)
    print(f"drop-in, deterministic: min {min(det):.4f} within 1e-12 over {len(det)} cases; "
          f"stochastic: min {min(rnd):.4f}, mean {np.mean(rnd):.5f} within 1e-9 over {len(rnd)} cases", flush=True)
    assert not bad, bad[:10]
    assert np.mean(rnd) >= 0.999


@pytest.mark.parametrize("random", [False, True])
def test_dropin_full_size_24h(random):
    """A 4,000-particle sample of the full-size release (every 1,000th of the
    4M, all 50 sources) moved 24 h through the drop-in entry point."""
    W = full_window()
    p = release(W, 4_000_000, 50)[::1000].copy()
    cfg = wc.default_config(random=random, sigma=wc.sigma_from_kh(1.0, 30.0))
    ref = run_oracle(W, cfg, p, 24)
    dev = run_dropin(W, cfg, p, 24)
    tol = 1e-9 if random else 1e-12
    ok = np.all(np.stack([close(dev[f], ref[f], tol) for f in ("k", "j", "i", "age")]), axis=0)
    worst = max(float(np.nanmax(np.abs(dev[f] - ref[f]))) for f in ("k", "j", "i"))
    print(f"full size 24 h, diffusion {'on' if random else 'off'}: {ok.mean():.5f} within {tol}, "
          f"largest |delta| {worst:.3e}", flush=True)
    assert ok.mean() >= (0.999 if random else 1.0)


def test_dropin_concentration_at_scale():
    """wc_hip_concentration (Concentration::evaluate, Concentration.hpp:16-44)
    on 1M particle_data records spread over the full-size frame -- .5 ties in
    k and j, dead particles, positions outside the grid -- against the
    oracle's counts, and adding into a frame that already holds counts (the
    reference adds +1.0f per particle into the frame it is given)."""
    import pyoracle as po
    from hip_rt import DeviceBuffer
    rng = np.random.default_rng(31)
    s_rho, eta, xi, n = 11, 1024, 768, 1_000_000
    p = wc.make_particles(n)
    p["k"] = rng.uniform(-s_rho + 0.4, 0.4, n)
    p["j"] = rng.uniform(-0.6, eta - 0.4, n)
    p["i"] = rng.uniform(-0.6, xi - 0.4, n)
    tie = rng.random(n) < 0.05
    p["j"][tie] = np.floor(p["j"][tie]) + 0.5
    p["k"][tie] = np.round(p["k"][tie]) - 0.5
    p["health"] = np.where(rng.random(n) < 0.1, -1.0, 1.0)
    # a clustered release on top: 200k particles in a few cells
    q = wc.make_particles(200_000, k=-0.2, j=500.3, i=400.7)
    q["j"] += np.repeat(np.arange(4), 50_000)
    p = np.concatenate([p, q]).astype(wc.PARTICLE_DTYPE)        # concatenate packs the records
    want = po.grid_count(p, s_rho, eta, xi).astype(np.float32)
    pd = DeviceBuffer(p)
    start = np.zeros((s_rho, eta, xi), np.float32)
    start[0, :2, :3] = 5.0
    cd = DeviceBuffer(start)
    L = wc.load_library()
    try:
        assert L.wc_hip_concentration(pd.ptr, len(p), s_rho, eta, xi, cd.ptr, None) == wc.WC_OK
        got = cd.get(start)
    finally:
        pd.free()
        cd.free()
    assert np.array_equal(got, want + start)
    assert float(want.sum()) > 0.8 * n

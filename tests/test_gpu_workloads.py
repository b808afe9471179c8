"""GPU tests at the stated workloads of the two north-star components that
have no reference counterpart:

Batched ensemble, the full 256-scenario workload (about 16M particles,
lognormal counts of median 40k in [10k, 500k], own source, Kh log-uniform in
[0.1, 10] m2/s, release offsets 0-6 h, BF16 forcing, 24 h, one captured
step replayed per interval):
  - every scenario's surface counts after 24 h equal a brute-force count of
    its alive, released, in-domain particles at rounded level 0;
  - a permuted population gives the same per-id result bit for bit (draws
    are keyed by id);
  - after the first interval, particles released later are unmoved and
    uncounted;
  - the particles of one scenario alone, moved on the mixed-precision
    context path with that scenario's sigma, end bit-identical (k, j, i,
    age; health within its FP32 storage).

Reconstruction network at the sizes it is trained and used at: the
forward pass and gradients on the training batch (16 x 128 x 128, three
input channels: histogram, log10 Kh, land mask) and one forward pass over
the whole 1024 x 768 surface, against the FP64 oracle with BF16 rounding
emulated, to the bars of tests/test_gpu_unet.py.
"""
import numpy as np
import pytest

import unet_oracle as uo
import wacommplusplus_amd as wc
from runner import run_device
from test_gpu_ensemble import brute_surface, load_window
from wacommplusplus_amd.synthetic import ensemble, full_window

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full():
    return full_window()


@pytest.fixture(scope="module")
def workload(full):
    return ensemble(full)


def run24(ctx, W, cfg, n=24):
    times = [float(W.times[q]) for q in range(n + 1)]
    return ctx.ens_run(cfg, times, list(range(n + 1)), use_graph=True, surface=True)


def test_ensemble_full_workload(full, workload):
    p, scen, sigma, counts = workload
    assert len(sigma) == 256 and 15_000_000 < len(p) < 17_000_000
    cfg = wc.default_config(random=True, sigma=1.0)
    with wc.Context(0) as ctx:
        load_window(ctx, full)
        ctx.ens_set(p, scen, sigma)
        initial, _ = ctx.ens_get()
        # first interval: the particles released after it stay put and uncounted
        run24(ctx, full, cfg, 1)
        one, sc1 = ctx.ens_get()
        late = p["time"] > full.times[1]
        assert 0.5 < late.mean() < 0.95
        for f in ("k", "j", "i", "age", "health"):
            assert np.array_equal(one[f][late], initial[f][late], equal_nan=True), f
        surf1 = ctx.ens_surface()
        assert np.array_equal(surf1, brute_surface(one, sc1, 256, full.eta, full.xi, full.times[1]))
        # the whole 24 h
        ctx.ens_set(p, scen, sigma)
        st = run24(ctx, full, cfg)
        out, sc = ctx.ens_get()
        surf = ctx.ens_surface()
        want = brute_surface(out, sc, 256, full.eta, full.xi, full.times[24])
        assert np.array_equal(surf, want)
        per = surf.reshape(256, -1).sum(1)
        # some releases leave the surface layer or die at the coast within 24 h
        assert (per > 0).sum() >= 200 and int(per.sum()) > 0.1 * len(p)
        print(f"ensemble 24 h: {len(p)} particles, {st.particle_steps:.3e} particle-steps in {st.device_ms:.0f} ms; "
              f"surface counts {int(per.sum())}", flush=True)
        # a permuted population: the same per-id result
        perm = np.random.default_rng(3).permutation(len(p))
        ctx.ens_set(p[perm], scen[perm], sigma)
        run24(ctx, full, cfg)
        outp, scp = ctx.ens_get()
        assert np.array_equal(scp, sc)
        for f in ("id", "k", "j", "i", "health", "age", "time"):
            assert np.array_equal(outp[f], out[f], equal_nan=f != "id"), f
        # one scenario alone on the context path, with its sigma
        s = int(np.argmax(counts < 20_000))
        m = scen == s
        ref = run_device(ctx, full, cfg.copy(sigma=float(sigma[s])), p[m], 24, path="fast", store="bf16")
    got = out[m]
    assert np.array_equal(got["id"], ref["id"])
    for f in ("k", "j", "i", "age"):
        assert np.array_equal(got[f], ref[f], equal_nan=True), f
    alive = ref["health"] > 0
    assert np.array_equal(got["health"] > 0, alive)
    assert np.max(np.abs(got["health"][alive] - ref["health"][alive])) < 1e-7


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def planes(x, kh, mask):
    """Histogram channel, constant log10 Kh plane, land mask -- the three
    input channels the reconstruction trains with."""
    n, h, w = x.shape
    return np.ascontiguousarray(np.stack([x, np.broadcast_to(np.log10(kh)[:, None, None], (n, h, w)),
                                          np.broadcast_to(mask, (n, h, w))], 1).astype(np.float32))


def test_unet_training_batch_matches_oracle(full):
    rng = np.random.default_rng(77)
    shape = (16, 128, 128)
    p = uo.he_params(seed=8, channels=3).astype(np.float32)
    x = planes(rng.gamma(0.5, 1.0, shape), rng.uniform(0.1, 10.0, 16), full.mask[:128, :128])
    t = (x[:, 0] + rng.normal(0.0, 0.2, shape)).astype(np.float32)
    net = wc.UNet(0, seed=1, channels=3)
    try:
        net.set_params(p)
        y, ms = net.forward(x)
        loss, g = net.gradients(x, t)
    finally:
        net.close()
    ora = uo.UNetOracle(p.astype(np.float64), emulate_bf16=True)
    y_ref = ora.forward(x)
    loss_ref, g_ref = ora.gradients(x, t)
    print(f"16x3x128x128: forward rel {rel(y, y_ref):.2e} ({ms:.2f} ms), loss {loss:.5g} vs {loss_ref:.5g}, "
          f"gradient rel {rel(g, g_ref):.2e}", flush=True)
    assert rel(y, y_ref) < 5e-3
    assert abs(loss - loss_ref) / loss_ref < 1e-3
    off = 0
    for l, (co, c1, c2) in enumerate(uo.SPEC):
        nw = co * 9 * (c1 + c2)
        ew, eb = rel(g[off:off + nw], g_ref[off:off + nw]), rel(g[off + nw:off + nw + co], g_ref[off + nw:off + nw + co])
        assert ew < 1e-1 and eb < 1e-1, (l, ew, eb)
        off += nw + co
    assert rel(g, g_ref) < 3e-2


def test_unet_full_surface_forward_matches_oracle(full):
    rng = np.random.default_rng(78)
    p = uo.he_params(seed=9, channels=3).astype(np.float32)
    x = planes(rng.gamma(0.5, 1.0, (1, full.eta, full.xi)), np.array([1.0]), full.mask)
    net = wc.UNet(0, seed=1, channels=3)
    try:
        net.set_params(p)
        y, ms = net.forward(x)
    finally:
        net.close()
    y_ref = uo.UNetOracle(p.astype(np.float64), emulate_bf16=True).forward(x)
    print(f"1x3x1024x768: forward rel {rel(y, y_ref):.2e} ({ms:.2f} ms)", flush=True)
    assert y.shape == (1, full.eta, full.xi)
    assert rel(y, y_ref) < 5e-3

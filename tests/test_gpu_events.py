"""GPU tests of closure-event tracking (wc_set_event_tracking / wc_get_events):
the per-particle WC_EVENT_* bits both move paths OR in when a closure acts.
They define the coast-pinned population of the precision comparison, so the
bits are checked against closed forms and against each other:

  - plain configuration (uniform u = 0.2 m/s, 200 m cells, no land, k fixed):
    no closure acts in the first 12 h; the plume reaches the east boundary
    (i = 99) at about 13.75 h, so after 24 h every particle carries
    WC_EVENT_DOMAIN, and at most the midpoint's fallback (its half step
    leaves the domain first), the same on both paths;
  - with tracking off nothing is written; set_particles and clear_events
    start from zero; compaction keeps each survivor's bits;
  - on the full-size window (land, w, coasts) the two paths flag the same
    particles for at least 99 % of a 20,000-particle sample over 4 h.
"""
import numpy as np
import pytest

import wacommplusplus_amd as wc
from runner import run_device
from wacommplusplus_amd.synthetic import full_window, plain_window, release

pytestmark = pytest.mark.gpu

EV = wc.EVENT


@pytest.fixture(scope="module")
def ctx():
    c = wc.Context(0)
    yield c
    c.close()


def tracked(ctx, W, cfg, p, n, **kw):
    ctx.set_event_tracking(True)
    try:
        out = run_device(ctx, W, cfg, p, n, **kw)
        return out, ctx.get_events()
    finally:
        ctx.set_event_tracking(False)


@pytest.mark.parametrize("path,store", [("parity", "f32"), ("fast", "f32"), ("fast", "bf16_tile")])
def test_events_plain_configuration(ctx, path, store):
    W = plain_window(T=25)
    p = release(W, 2000, 1, depth_m=1.0)
    p["j"], p["i"] = 59.5, 49.5
    cfg = wc.default_config(random=False)
    _, ev12 = tracked(ctx, W, cfg, p, 12, path=path, store=store)
    assert np.all(ev12 == 0), np.unique(ev12)
    out, ev24 = tracked(ctx, W, cfg, p, 24, path=path, store=store)
    assert np.all(ev24 & EV["domain"]), np.unique(ev24)
    allowed = EV["domain"] | EV["fallback"]
    assert np.all((ev24 | allowed) == allowed), np.unique(ev24)
    assert len(np.unique(ev24)) == 1                       # one trajectory: one outcome
    assert np.all(out["health"] > 0) and np.all(out["i"] < 99)     # reflected back inside


def test_events_off_clear_and_compaction(ctx):
    W = plain_window(T=25)
    p = release(W, 3000, 1, depth_m=1.0)
    p["j"], p["i"] = 59.5, 49.5
    cfg = wc.default_config(random=False)
    run_device(ctx, W, cfg, p, 24)                          # tracking off
    assert np.all(ctx.get_events() == 0)
    # kill at the open boundary: the particles released further east die there
    kill = cfg.copy(horizontalClosure=wc.KILL)
    q = p.copy()
    q["i"] = np.linspace(40.0, 60.0, len(q))
    out, ev = tracked(ctx, W, kill, q, 14)
    alive = out["health"] > 0
    assert 0 < alive.sum() < len(q)
    assert np.all(ev[~alive] & EV["domain"]) and np.all(ev[alive] == 0)
    ctx.remove_dead()
    assert np.array_equal(ctx.get_events(), ev[alive])
    ctx.clear_events()
    assert np.all(ctx.get_events() == 0)
    ctx.set_particles(q)
    assert np.all(ctx.get_events() == 0)
    ctx.set_event_tracking(False)


def test_events_agree_between_paths(ctx):
    W = full_window(T=5)
    p = release(W, 4_000_000, 50, spread=0.25)[::200].copy()
    cfg = wc.default_config(random=False)
    a, ev_a = tracked(ctx, W, cfg, p, 4, path="parity")
    b, ev_b = tracked(ctx, W, cfg, p, 4, path="fast", store="f32")
    pin = EV["shore"] | EV["domain"] | EV["fallback"]
    same = ((ev_a & pin) != 0) == ((ev_b & pin) != 0)
    print(f"pinned on FP64 {np.mean((ev_a & pin) != 0):.3f}, FP32 {np.mean((ev_b & pin) != 0):.3f}, "
          f"agree {same.mean():.4f}; coastal {np.mean((ev_a & EV['coast']) != 0):.3f}", flush=True)
    assert same.mean() >= 0.99
    assert np.mean(((ev_a & EV["coast"]) != 0) == ((ev_b & EV["coast"]) != 0)) >= 0.99

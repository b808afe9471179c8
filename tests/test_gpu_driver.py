"""GPU tests of the release (wc_emit) and of the interval driver
(wc_run_window) against the oracle: Sources::emit / Source::emit /
sourceEmissionTimes (Sources.cpp:233-248, Source.cpp:25-44,
SourceEmission.hpp:38-79) and the interval loop of Wacomm::run
(Wacomm.cpp:200-770: emit -> move -> remove dead -> grid)."""
import numpy as np
import pytest

import pyoracle as po
import wacommplusplus_amd as wc
from runner import cfg_to_oracle, oracle_pair
from wacommplusplus_amd.synthetic import full_window

pytestmark = pytest.mark.gpu

SOURCE_POSITION_RANDOM_INTERVAL = -(2**63) + 2          # SourceEmission.hpp:18


def oracle_emit(cfg, sources, start, end, particles):
    """Sources::emit restated with the oracle's emission times and keyed normals."""
    L = po.lib()
    em = []
    for q, s in enumerate(sources):
        t = po.emission_times(s.mode, s.value, s.start, s.end, start, end)
        assert not isinstance(t, int)
        em += [(x, q) for x in t]
    em.sort(key=lambda e: e[0])                          # stable: equal times keep source order
    out = []
    next_id = 0 if len(particles) == 0 else int(particles["id"][-1]) + 1
    for x, q in em:
        s = sources[q]
        k, j, i = s.k, s.j, s.i
        if cfg.randomSources:
            k = k + 0.25 * L.wco_normal(cfg.randomSeed, next_id, SOURCE_POSITION_RANDOM_INTERVAL, 0, 0)
            j = j + 0.25 * L.wco_normal(cfg.randomSeed, next_id, SOURCE_POSITION_RANDOM_INTERVAL, 0, 1)
            i = i + 0.25 * L.wco_normal(cfg.randomSeed, next_id, SOURCE_POSITION_RANDOM_INTERVAL, 0, 2)
        side = cfg.driftSide                              # Source.cpp:38-42
        if cfg.leewayRandomSide:
            side = L.wco_sample_drift_side(cfg.randomSeed, next_id, cfg.leewayRightSideProbability)
        out.append((next_id, k, j, i, x, side))
        next_id += 1
    res = np.zeros(len(particles) + len(out), dtype=wc.PARTICLE_DTYPE)   # keeps the 64-byte layout
    res[: len(particles)] = particles
    for q, (pid, k, j, i, x, side) in enumerate(out):
        r = res[len(particles) + q]
        r["id"], r["k"], r["j"], r["i"], r["health"], r["age"], r["time"] = pid, k, j, i, 1.0, 0.0, x
        r["driftObjectType"], r["driftSide"] = cfg.driftObjectType, side
    return res


@pytest.fixture(scope="module")
def small():
    return full_window(T=7, eta=96, xi=80, s_rho=6, spacing=150.0, seed=17)


def sources_for(W):
    return [wc.Source(-0.3, 40.3, 30.7, wc.UNIFORM_RATE, 1000.0, 0.0, -1.0),
            wc.Source(0.0, 50.5, 45.2, wc.SINGLE_PULSE, 300.0, 1800.0, -1.0),
            wc.Source(-1.0, 60.1, 20.9, wc.FORCING_INTERVAL_BATCH, 120.0, 3600.0, 10800.0),
            wc.Source(-0.5, 30.2, 50.4, wc.UNIFORM_RATE, 77.0, 900.0, 5000.0)]


DRIFT = {"passive": {}, "typed": dict(driftObjectType=3, driftSide=-1),
         "random_side": dict(driftObjectType=6, leewayRandomSide=True, leewayRightSideProbability=0.3)}


@pytest.mark.parametrize("drift", list(DRIFT))
@pytest.mark.parametrize("random_sources", [False, True])
def test_emit_matches_oracle(small, random_sources, drift):
    W = small
    cfg = wc.default_config(randomSources=random_sources, randomSeed=42, **DRIFT[drift])
    src = sources_for(W)
    with wc.Context(0) as ctx:
        ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask)
        ref = wc.make_particles(0)
        for t in range(4):
            ctx.emit(cfg, src, W.times[t], W.times[t + 1])
            ref = oracle_emit(cfg, src, W.times[t], W.times[t + 1], ref)
        dev = ctx.get_particles()
    assert len(dev) == len(ref) > 0
    assert np.array_equal(dev["id"], ref["id"]) and np.array_equal(dev["time"], ref["time"])
    assert np.array_equal(dev["driftObjectType"], ref["driftObjectType"])
    assert np.array_equal(dev["driftSide"], ref["driftSide"])
    if drift == "random_side":
        assert set(np.unique(dev["driftSide"])) == {-1, 1}
    tol = 1e-12 if random_sources else 0.0            # device log/cos vs glibc: ulps of 0.25*normal
    for f in ("k", "j", "i"):
        assert np.max(np.abs(dev[f] - ref[f])) <= tol


@pytest.mark.parametrize("n_slots", [2, 3, 7])
def test_run_window_matches_oracle(small, n_slots):
    """wc_run_window (emit, move, remove dead, grid per interval) against the
    same loop driven on the oracle; deterministic, so bit-exact.  Two slots
    alternate uploads and moves; three or more upload the record after next
    on the copy stream while the move runs."""
    W = small
    cfg = wc.default_config(random=False, randomSources=False)
    src = sources_for(W)
    zeta, u, v, w, a = W.stacked()
    with wc.Context(0) as ctx:
        ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask, n_slots=n_slots)
        conc, st = ctx.run_window(cfg, W.times, zeta, u, v, w, a, sources=src, mask_output=True)
        dev = ctx.get_particles()
    assert st.intervals == W.T - 1
    ref = wc.make_particles(0)
    ocfg = cfg_to_oracle(cfg)
    land = W.mask != 1
    for t in range(W.T - 1):
        ref = oracle_emit(cfg, src, W.times[t], W.times[t + 1], ref)
        win, idx = oracle_pair(W, t, t + 1)
        po.move(win, ocfg, idx, ref)
        ref = np.ascontiguousarray(ref[ref["health"] > 0])
        frame = po.grid_count(ref, W.s_rho, W.eta, W.xi).astype(np.float32)
        frame[:, land] = np.float32(1e37)
        assert np.array_equal(conc[t + 1], frame), t
    assert np.array_equal(conc[0], np.zeros_like(conc[0]))
    assert len(dev) == len(ref) == st.particles
    assert np.array_equal(dev["id"], ref["id"])
    for f in ("k", "j", "i", "age", "time"):
        assert np.array_equal(dev[f], ref[f], equal_nan=True), f

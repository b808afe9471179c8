"""GPU parity of libwacomm_hip.so against the CPU oracle (the reference's
Particle::move / Concentration::evaluate restated in oracle/wc_oracle.c).

Bars (tests/CudaParticleParityTest.cu:141 and tools/performance_evidence.py:11):
  - WC_PATH_PARITY, diffusion off: positions, age and health bit-identical
    (the kernel keeps every FP64 rounding point of the CPU path; only exp()
    of the decay can differ, by <= 1 ulp);
  - WC_PATH_PARITY, diffusion on: |delta| <= 1e-9 cells (device log/cos/sqrt
    may differ from glibc by an ulp);
  - gridding: counts bit-exact on identical positions.
Every test calls the library through its C ABI.
"""
import ctypes as C
import math

import numpy as np
import pytest

import pyoracle as po
import wacommplusplus_amd as wc
from runner import max_cell_error, oracle_pair, run_device, run_oracle, cfg_to_oracle
from scenarios import DEG, tiny_window
from wacommplusplus_amd.synthetic import ForcingWindow, full_window, plain_window, release

pytestmark = pytest.mark.gpu

NAN = float("nan")
# closure events of the device path (WC_EVENT_*), the coast-cell bit aside
CLOSURE_EVENTS = (wc.EVENT["shore"] | wc.EVENT["domain"] | wc.EVENT["fallback"] | wc.EVENT["vertical"] |
                  wc.EVENT["shallow"])


@pytest.fixture(scope="module")
def ctx():
    c = wc.Context(0)
    yield c
    c.close()


def as_window(tw, s_rho=2):
    """tests/scenarios.py tiny window (oracle Window) -> synthetic ForcingWindow."""
    recs = [(tw.zeta[t], tw.u[t], tw.v[t], tw.w[t], tw.akt[t]) for t in range(tw.T)]
    return ForcingWindow(tw.s_rho, tw.eta, tw.xi, tw.lonRad, tw.latRad, tw.h,
                         np.ones((tw.eta, tw.xi)), tw.depthIntervals[: tw.s_w - 1], tw.ocean_time,
                         lambda t: recs[t])


def move_cfg(**kw):
    base = dict(random=False, deltat=3600, dti=30, survprob=0, tau0=86400, crid=1, sigma=3.46,
                shoreLimit=.25, upperClosure=1, lowerClosure=1, horizontalClosure=1, trackingDirection=1,
                restartCheckpoint=NAN, integrator=0, randomSeed=5489)
    base.update(kw)
    return wc.default_config(**base)


def one(id_, k, j, i, time, health=1.0, age=0.0):
    p = wc.make_particles(1, k=k, j=j, i=i, time=time, health=health, age=age)
    p["id"] = id_
    return p


def close(a, b, tol):
    """|a-b| <= tol elementwise, a NaN on both sides counting as equal (the
    reference itself produces NaN k where a particle sits on an all-land cell:
    zdist = 0 in Particle.cpp:579, 657)."""
    return (np.abs(a - b) <= tol) | (np.isnan(a) & np.isnan(b))


def assert_same(dev, ref, tol=0.0, fields=("k", "j", "i", "age")):
    assert len(dev) == len(ref)
    assert np.array_equal(dev["id"], ref["id"])
    for f in fields:
        ok = close(dev[f], ref[f], tol)
        bad = np.where(~ok)[0]
        assert ok.all(), (f, bad[:5], dev[f][bad[:5]], ref[f][bad[:5]])
    assert close(dev["health"], ref["health"], 1e-15).all()


# --------------------------------------------------------------------------
# the reference's own move cases (tests/ParticleMoveTest.cpp) on the 2x2 grid
# --------------------------------------------------------------------------

U_RAMP = np.zeros((2, 2, 2, 2), np.float32)
U_RAMP[1] = 2.0
SHEAR = np.zeros((2, 2, 2, 2), np.float32)
SHEAR[:, :, :, 1] = 10.0
W_SURF = np.zeros((2, 3, 2, 2), np.float32)
W_SURF[:, 2] = 2.0

CASES = [
    # name, window kwargs, config kwargs, particle, direction
    ("time_linear", dict(u=U_RAMP), {}, (1, -.5, .25, .25, 0)),
    ("late_emission", dict(u=U_RAMP), {}, (2, -.5, .25, .25, 40)),
    ("backward", dict(u=U_RAMP), dict(trackingDirection=-1), (3, -.5, .25, .2505846, 65)),
    ("stochastic", dict(u=U_RAMP), dict(random=True), (4, -.5, .25, .25, 0)),
    ("vertical_w", dict(times=(0.0, 1.0), w=W_SURF), {}, (7, -.5, .25, .25, 0)),
    ("vertical_w_025", dict(times=(0.0, 1.0), w=W_SURF), {}, (7, -.25, .25, .25, 0)),
    ("shear_euler", dict(times=(0.0, 3600.0), u=SHEAR), dict(integrator=1, dti=300), (40, -.5, .25, .25, 0)),
    ("shear_midpoint", dict(times=(0.0, 3600.0), u=SHEAR), dict(integrator=2, dti=300), (41, -.5, .25, .25, 0)),
    ("midpoint_fallback", dict(times=(0.0, 600.0), u=SHEAR), dict(integrator=2, dti=600), (43, -.5, .25, .99, 0)),
    ("restart_checkpoint", dict(u=U_RAMP), dict(restartCheckpoint=30.0, random=True), (8, -.5, .25, .25, 0)),
    ("upper_reflect", {}, dict(upperClosure=3), (11, .3, .25, .25, 0)),
    ("lower_reflect", {}, dict(lowerClosure=3), (12, -1.2, .25, .25, 0)),
    ("lower_kill", {}, dict(lowerClosure=2), (12, -1.2, .25, .25, 0)),
    ("open_boundary", {}, {}, (13, -.5, 1.2, .25, 0)),
    ("decay", {}, dict(tau0=100.0), (9, -.5, .25, .25, 0)),
    ("survival", {}, dict(survprob=0.6), (10, -.5, .25, .25, 0, 0.5)),
]


@pytest.mark.parametrize("name,wkw,ckw,part", CASES, ids=[c[0] for c in CASES])
def test_reference_move_cases_context(ctx, name, wkw, ckw, part):
    tw = tiny_window(**wkw)
    W = as_window(tw)
    cfg = move_cfg(**ckw)
    p = one(*part)
    ref = p.copy()
    t = 0 if cfg.trackingDirection > 0 else 1
    po.move(tw, cfg_to_oracle(cfg), t, ref)
    dev = run_device(ctx, W, cfg, p, 1, first=t)
    assert_same(dev, ref, tol=1e-9 if cfg.random else 0.0)


def test_reference_parity_modes_dropin():
    """tests/CudaParticleParityTest.cu:30-158 through the drop-in entry point
    (device pointers, particle_data AoS): passive modes 0 (deterministic),
    1 (stochastic), 9 (backward), 10 (restart at 35 s) under Euler and
    midpoint; id 2^53+1; tolerances 1e-12 / 1e-9."""
    from hip_rt import DeviceBuffer
    u = np.zeros(16, np.float32)
    u[8:] = .25
    akt = np.full(24, .01, np.float32)
    host = dict(ocean=np.array([0.0, 65.0]), lon=np.array([0, DEG, 0, DEG]), lat=np.array([0, 0, DEG, DEG]),
                depth=np.array([.5, .5, .5]), h=np.full(4, 100.0), mask=np.ones(4),
                zeta=np.zeros(8, np.float32), u=u, v=np.zeros(16, np.float32), w=np.zeros(24, np.float32),
                akt=akt)
    d = {k: DeviceBuffer(v) for k, v in host.items()}
    tw = tiny_window(u=u.reshape(2, 2, 2, 2), akt=akt.reshape(2, 3, 2, 2))
    L = wc.load_library()
    try:
        for run in range(22):
            mode = run % 11
            if mode not in (0, 1, 9, 10):
                continue
            cfg = move_cfg(random=mode == 1, sigma=.1, integrator=1 if run < 11 else 2,
                           trackingDirection=-1 if mode >= 9 else 1,
                           restartCheckpoint=35.0 if mode == 10 else NAN)
            time_index = 1 if mode >= 9 else 0
            p = one(9007199254740993, -.5, .25, .25, 0)
            ref = p.copy()
            po.move(tw, cfg_to_oracle(cfg), time_index, ref)
            cfg_dev = DeviceBuffer(np.frombuffer(bytes(cfg), np.uint8))
            p_dev = DeviceBuffer(p)
            rc = L.wc_hip_move_particles(cfg_dev.ptr, p_dev.ptr, time_index, 2, 3, 2, 2, 2, d["ocean"].ptr,
                                         d["mask"].ptr, d["lon"].ptr, d["lat"].ptr, d["depth"].ptr, d["h"].ptr,
                                         d["zeta"].ptr, d["u"].ptr, d["v"].ptr, d["w"].ptr, d["akt"].ptr,
                                         None, None, None, None, 1, 1, 1, None)
            assert rc == 0
            out = p_dev.get(p)
            cfg_dev.free()
            p_dev.free()
            tol = 1e-9 if mode == 1 else 1e-12
            for f in ("i", "j", "k", "health", "age", "time"):
                assert abs(out[f][0] - ref[f][0]) <= tol, (mode, run, f, out[f][0], ref[f][0])
            assert out["id"][0] == ref["id"][0]
    finally:
        for b in d.values():
            b.free()


def test_unsupported_drift_model(ctx):
    W = as_window(tiny_window())
    cfg = move_cfg(driftModel=1)
    from runner import load_grid
    load_grid(ctx, W)
    ctx.upload_record(0, *W.record(0))
    ctx.upload_record(1, *W.record(1))
    ctx.set_particles(one(1, -.5, .25, .25, 0))
    with pytest.raises(wc.WcError) as e:
        ctx.move_interval(cfg, 0.0, 65.0, 0, 1)
    assert e.value.status == 3


# --------------------------------------------------------------------------
# randomised small windows with land: every closure, both integrators
# --------------------------------------------------------------------------

def random_window(seed, eta=24, xi=20, s_rho=4, T=4):
    """A small window with land, free-surface elevation, strong currents and a
    large AKt, so that every closure fires within a few hours.  The velocity
    gradients stay moderate (a difference grows by well under 1 % per substep):
    the stochastic comparison below is then a statement about the draws and
    not about the amplification of ulps by a chaotic field."""
    rng = np.random.default_rng(seed)
    W = full_window(T=T, eta=eta, xi=xi, s_rho=s_rho, spacing=400.0, seed=seed, n_modes=16, rms=1.5,
                    land_fraction=0.25, H_decay=15.0, akt=0.1)
    zetas = [rng.uniform(-0.3, 0.3, (eta, xi)).astype(np.float32) * (W.mask > 0) for _ in range(T)]
    base = W.record_fn

    def rec(t):
        z, u, v, w, a = base(t)
        return zetas[t], u, v, w * 2.0, a
    W.record_fn = rec
    W._cache.clear()
    return W


@pytest.mark.parametrize("seed", range(6))
def test_random_windows_all_closures(ctx, seed):
    rng = np.random.default_rng(100 + seed)
    W = random_window(seed)
    n = 400
    p = wc.make_particles(n)
    p["k"] = rng.uniform(-3.2, 0.2, n)
    p["j"] = rng.uniform(-0.5, W.eta - 0.5, n)
    p["i"] = rng.uniform(-0.5, W.xi - 0.5, n)
    p["time"] = rng.choice([0.0, 0.0, 900.0, 2000.0], n)
    p["id"] = rng.integers(0, 2**63, n, dtype=np.uint64)
    for integ in (1, 2):
        for closures in ((1, 1, 1), (2, 2, 2), (3, 3, 3), (1, 2, 3)):
            for random in (False, True):
                case = dict(integrator=integ, closures=closures, random=random, seed=seed)
                cfg = wc.default_config(integrator=integ, upperClosure=closures[0], lowerClosure=closures[1],
                                        horizontalClosure=closures[2], random=random, sigma=30.0,
                                        survprob=1e-4, tau0=20000.0, randomSeed=seed)
                ref = run_oracle(W, cfg, p, 3)
                ctx.set_event_tracking(True)
                dev = run_device(ctx, W, cfg, p, 3)
                ev = ctx.get_events()
                ctx.set_event_tracking(False)
                if not random:
                    assert_same(dev, ref, tol=0.0), case
                    continue
                ok = np.all(np.stack([close(dev[f], ref[f], 1e-9) for f in ("k", "j", "i")]), axis=0)
                # an ulp of the device log/cos can move a particle across a
                # closure threshold (shore limit, land, domain edge); this
                # window is built to hit closures constantly, so allow 3 %,
                # and every particle beyond 1e-9 took a closure branch
                print(f"{case}: {ok.mean():.4f} within 1e-9", flush=True)
                assert ok.mean() >= 0.97, (case, float(ok.mean()), np.where(~ok)[0][:5])
                assert np.all((ev[~ok] & CLOSURE_EVENTS) != 0), (case, np.where(~ok & (ev == 0))[0][:5])


@pytest.mark.parametrize("seed", range(2))
def test_random_windows_decay_threshold(ctx, seed):
    """Decay across the survival threshold inside an interval
    (Particle.cpp:218-223, 797-804): tau0 = 15000 s and survprob = 0.3 put
    the kill age at 18060 s, and the particles start at ages 0..12000 s, so
    in the 3 h window health is evaluated on both sides of it, particles die
    by decay at substeps throughout the window, and some are killed by a
    closure first.
    dti = 70 s leaves a shortened last step in every interval (a whole-step
    and a partial-step diffusion scale, a substep index that is not
    elapsed / dti of a whole step)."""
    rng = np.random.default_rng(300 + seed)
    W = random_window(seed)
    n = 400
    p = wc.make_particles(n)
    p["k"] = rng.uniform(-3.2, 0.2, n)
    p["j"] = rng.uniform(-0.5, W.eta - 0.5, n)
    p["i"] = rng.uniform(-0.5, W.xi - 0.5, n)
    p["time"] = rng.choice([0.0, 0.0, 900.0, 2000.0], n)
    p["age"] = rng.uniform(0.0, 12000.0, n)
    kill_age = -15000.0 * np.log(0.3)
    p["id"] = rng.integers(0, 2**63, n, dtype=np.uint64)
    for integ in (1, 2):
        for random in (False, True):
            case = dict(integrator=integ, random=random, seed=seed)
            cfg = wc.default_config(integrator=integ, upperClosure=3, lowerClosure=2, horizontalClosure=2,
                                    random=random, sigma=30.0, dti=70.0, survprob=0.3, tau0=15000.0,
                                    randomSeed=seed)
            ref = run_oracle(W, cfg, p, 3)
            ctx.set_event_tracking(True)
            dev = run_device(ctx, W, cfg, p, 3)
            ev = ctx.get_events()
            ctx.set_event_tracking(False)
            dead = ref["health"] < 0
            assert np.count_nonzero(dead & (ref["age"] > kill_age)) >= 20, case     # died by decay
            if not random:
                assert_same(dev, ref, tol=0.0), case
                continue
            ok = np.all(np.stack([close(dev[f], ref[f], 1e-9) for f in ("k", "j", "i")]), axis=0)
            ok &= close(dev["health"], ref["health"], 1e-15)
            print(f"{case}: {ok.mean():.4f} within 1e-9, {dead.mean():.3f} dead", flush=True)
            assert ok.mean() >= 0.97, (case, float(ok.mean()), np.where(~ok)[0][:5])
            assert np.all((ev[~ok] & CLOSURE_EVENTS) != 0), (case, np.where(~ok & (ev == 0))[0][:5])


@pytest.mark.parametrize("seed", range(3))
def test_random_windows_backward_and_restart(ctx, seed):
    """Backward tracking from the last record (Particle.cpp:81-125: reversed
    interval, direction -1; diffusion only with backwardDiffusion) and a
    restart checkpoint inside the first interval (restartElapsed,
    NumericalHelpers.hpp:45-52), on the land-heavy random windows."""
    rng = np.random.default_rng(200 + seed)
    W = random_window(seed)
    n = 400
    p = wc.make_particles(n)
    p["k"] = rng.uniform(-3.2, 0.2, n)
    p["j"] = rng.uniform(-0.5, W.eta - 0.5, n)
    p["i"] = rng.uniform(-0.5, W.xi - 0.5, n)
    p["time"] = rng.choice([0.0, 900.0, 2000.0, 1e5], n)
    p["id"] = rng.integers(0, 2**63, n, dtype=np.uint64)
    cases = []
    for integ in (1, 2):
        for random, bwd in ((False, 0), (True, 0), (True, 1)):
            cases.append(dict(integrator=integ, random=random, backwardDiffusion=bwd, trackingDirection=-1))
        for random in (False, True):
            cases.append(dict(integrator=integ, random=random, restartCheckpoint=float(W.times[0]) + 1000.0))
    for kw in cases:
        cfg = wc.default_config(sigma=30.0, survprob=1e-4, tau0=20000.0, randomSeed=seed, upperClosure=1,
                                lowerClosure=2, horizontalClosure=3, **kw)
        ref = run_oracle(W, cfg, p, 3)
        dev = run_device(ctx, W, cfg, p, 3)
        diffusing = kw["random"] and (kw.get("trackingDirection", 1) > 0 or kw.get("backwardDiffusion", 0))
        if not diffusing:
            print(f"{kw}: bit-identical check", flush=True)
            assert_same(dev, ref, tol=0.0)
            continue
        ok = np.all(np.stack([close(dev[f], ref[f], 1e-9) for f in ("k", "j", "i")]), axis=0)
        print(f"{kw}: {ok.mean():.4f} within 1e-9", flush=True)
        assert ok.mean() >= 0.97, (kw, float(ok.mean()))


# --------------------------------------------------------------------------
# the plain CPU configuration: 10,000 particles, 24 h
# --------------------------------------------------------------------------

def test_plain_configuration_24h(ctx):
    W = plain_window()
    p = release(W, 10000, 1, depth_m=1.0)
    p["j"], p["i"] = 59.5, 49.5
    cfg = wc.default_config(random=False)
    ref = run_oracle(W, cfg, p, 24)
    dev = run_device(ctx, W, cfg, p, 24)
    assert_same(dev, ref, tol=0.0)
    assert np.all(dev["k"] == -0.2)


# --------------------------------------------------------------------------
# the full-size configuration (1024 x 768 x 11, land, 64 modes): a sample of
# the 4M release followed for 24 h, diffusion off and on
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def full():
    return full_window()


@pytest.mark.parametrize("random", [False, True], ids=["diffusion_off", "diffusion_on"])
def test_full_configuration_sample_24h(ctx, full, random):
    p = release(full, 4_000_000, 50)[::1000].copy()          # 4,000 particles, all 50 sources
    cfg = wc.default_config(random=random, sigma=wc.sigma_from_kh(1.0, 30.0), crid=1.0)
    ref = run_oracle(full, cfg, p, 24)
    ctx.set_event_tracking(True)
    dev = run_device(ctx, full, cfg, p, 24)
    ev = ctx.get_events()
    ctx.set_event_tracking(False)
    if not random:
        assert_same(dev, ref, tol=0.0)
    else:
        ok = np.all(np.stack([close(dev[f], ref[f], 1e-9) for f in ("k", "j", "i")]), axis=0)
        # every particle within 1e-9 cells, except where an ulp difference of
        # the device log/cos flips a closure branch; at most 1 in 1000, and
        # each of those took a closure branch
        frac = float(np.mean(ok))
        print(f"diffusion on: {frac:.5f} of particles within 1e-9 cells", flush=True)
        assert frac >= 0.999, frac
        assert np.all((ev[~ok] & CLOSURE_EVENTS) != 0)


def test_upload_waits_for_the_move_reading_its_slot(ctx, full):
    """Uploads run on their own stream: an upload into a slot that an
    enqueued move still reads waits for that move (the per-slot event), and
    one into a free slot runs beside it.  2M particles of the full-size
    release keep the move (~13 ms) longer than a record's copy (~6 ms)."""
    p = release(full, 2_000_000, 50)
    cfg = wc.default_config(random=True, sigma=wc.sigma_from_kh(1.0, 30.0))
    ref = run_device(ctx, full, cfg, p, 1)
    from runner import load_grid
    load_grid(ctx, full, 3)
    ctx.upload_record(0, *full.record(0))
    ctx.upload_record(1, *full.record(1))
    ctx.set_particles(p)
    ctx.move_interval(cfg, float(full.times[0]), float(full.times[1]), 0, 1)     # enqueued, not waited for
    ctx.upload_record(1, *full.record(2))           # the slot the move reads: must wait for it
    ctx.upload_record(2, *full.record(3))           # a free slot: overlaps the move
    got = ctx.get_particles()
    for f in ref.dtype.names:
        assert np.array_equal(got[f], ref[f], equal_nan=f not in ("id", "driftObjectType", "driftSide")), f
    # the slots hold the new records
    z1 = ctx.get_record(1)[0]
    assert np.array_equal(z1, np.asarray(full.record(2)[0], np.float32))


# --------------------------------------------------------------------------
# gridding and dead removal
# --------------------------------------------------------------------------

def test_grid_counts_bit_exact(ctx):
    rng = np.random.default_rng(7)
    W = random_window(3, eta=30, xi=25, s_rho=5)
    n = 200000
    p = wc.make_particles(n)
    p["k"] = np.round(rng.uniform(-5.5, 0.6, n) * 2) / 2        # many exact .5 ties
    p["j"] = np.round(rng.uniform(-1.0, 31.0, n) * 4) / 4
    p["i"] = rng.uniform(-1.0, 26.0, n)
    p["health"] = rng.choice([1.0, 0.5, 0.0, -1.0], n)
    from runner import load_grid
    load_grid(ctx, W)
    ctx.set_particles(p)
    counts = ctx.grid()
    ref = po.grid_count(p, W.s_rho, W.eta, W.xi)
    assert np.array_equal(counts, ref)
    # the reference's float frame and the land mask (Wacomm.cpp:722-739)
    conc = ctx.grid(counts=False, conc=True, mask_output=True)
    land = W.mask == 0
    assert np.all(conc[:, land] == np.float32(1e37))
    assert np.array_equal(conc[:, ~land], ref[:, ~land].astype(np.float32))


def test_concentration_reference_case_dropin():
    """tests/ConcentrationTest.cpp through wc_hip_concentration."""
    from hip_rt import DeviceBuffer
    n = 100000
    p = wc.make_particles(n + 2, k=-1, j=1, i=1)
    p["j"][n] = 4
    p["health"][n + 1] = 0
    pd = DeviceBuffer(p)
    zero = np.zeros((3, 3, 3), np.float32)
    conc = DeviceBuffer(zero)
    L = wc.load_library()
    assert L.wc_hip_concentration(pd.ptr, n + 2, 3, 3, 3, conc.ptr, None) == 0
    c = conc.get(zero)
    pd.free()
    conc.free()
    assert c[1, 1, 1] == n and c.sum() == n


def test_remove_dead_order_preserving(ctx):
    rng = np.random.default_rng(11)
    W = as_window(tiny_window())
    from runner import load_grid
    load_grid(ctx, W)
    for n in (1, 255, 256, 257, 100003, 1 << 20):
        p = wc.make_particles(n, k=-.5, j=.5, i=.5)
        p["health"] = np.where(rng.random(n) < 0.3, -1.0, rng.random(n) + 0.01)
        ctx.set_particles(p)
        assert ctx.remove_dead() == int((p["health"] > 0).sum())
        assert np.array_equal(ctx.get_particles(), p[p["health"] > 0])


# --------------------------------------------------------------------------
# invariants at full size (4M particles)
# --------------------------------------------------------------------------

def test_full_size_invariants(ctx, full):
    p = release(full, 4_000_000, 50)
    cfg = wc.default_config(random=True, sigma=wc.sigma_from_kh(1.0, 30.0))
    a, frames = run_device(ctx, full, cfg, p, 2, grid_frames=True)
    # gridding conserves particles: every alive in-domain particle counted once
    alive = a[a["health"] > 0]
    kr, jr, ir = np.round(alive["k"]), np.round(alive["j"]), np.round(alive["i"])
    inside = (jr >= 0) & (jr < full.eta) & (ir >= 0) & (ir < full.xi) & (kr >= -full.s_rho + 1) & (kr <= 0)
    assert int(frames[-1].sum()) == int(inside.sum())
    # schedule independence: a permuted population gives the same per-id results
    perm = np.random.default_rng(5).permutation(len(p))
    b = run_device(ctx, full, cfg, p[perm], 2)
    order = np.argsort(b["id"])
    assert np.array_equal(b[order], a)

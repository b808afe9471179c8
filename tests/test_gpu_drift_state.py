"""GPU tests of the drift state a particle carries (particle_data's
driftObjectType and driftSide, Particle.hpp:31-41) through every device
path, against the oracle, through the C ABI.

With driftModel 0 a non-passive particle moves exactly as a passive one (the
leeway terms of Particle.cpp:380-454 act only under driftModel 1); its type is
carried unchanged, and with leewayJibeProbabilityHourly > 0 its side flips
after a live substep when the keyed uniform of component 'JIBE' falls below
jibeStepProbability (Particle.cpp:805-813, DriftModel.hpp:156-160).

Bars: types and sides bit-identical to the oracle; positions to the bars of
the parity tests (bit-identical without diffusion, 1e-9 cells with it).
"""
import numpy as np
import pytest

import pyoracle as po
import wacommplusplus_amd as wc
from runner import cfg_to_oracle, run_device, run_oracle
from scenarios import DEG, tiny_window
from test_gpu_parity import as_window, close, move_cfg, random_window

pytestmark = pytest.mark.gpu

PERSON_IN_WATER, RIGHT, LEFT = 1, 1, -1


@pytest.fixture(scope="module")
def ctx():
    c = wc.Context(0)
    yield c
    c.close()


def mixed_population(W, n, seed):
    """Positions over the window, a third of them passive, the rest of the 12
    drift object types (DriftModel.hpp:14-27) with sides -1, 0 and +1."""
    rng = np.random.default_rng(seed)
    p = wc.make_particles(n)
    p["k"] = rng.uniform(-2.5, -0.1, n)
    p["j"] = rng.uniform(1.0, W.eta - 2.0, n)
    p["i"] = rng.uniform(1.0, W.xi - 2.0, n)
    p["id"] = rng.integers(0, 2**63, n, dtype=np.uint64)
    p["driftObjectType"] = np.where(rng.random(n) < 1 / 3, 0, rng.integers(1, 12, n))
    p["driftSide"] = rng.choice([-1, 0, 1], n)
    return p


@pytest.mark.parametrize("integrator", [1, 2])
@pytest.mark.parametrize("random", [False, True])
def test_jibe_context_parity_path(ctx, integrator, random):
    W = random_window(3)
    p = mixed_population(W, 2000, seed=integrator)
    cfg = wc.default_config(integrator=integrator, random=random, sigma=30.0, upperClosure=1, lowerClosure=1,
                            horizontalClosure=1, survprob=0.0, tau0=20000.0, randomSeed=8,
                            leewayJibeProbabilityHourly=0.6)
    ref = run_oracle(W, cfg, p, 3)
    dev = run_device(ctx, W, cfg, p, 3, path="parity")
    print(f"integrator {integrator} random {random}: {np.mean(dev['driftSide'] != p['driftSide']):.3f} flipped",
          flush=True)
    assert np.array_equal(dev["id"], ref["id"])
    assert np.array_equal(dev["driftObjectType"], p["driftObjectType"])
    flipped = dev["driftSide"] != p["driftSide"]
    assert flipped.any() and not flipped[(p["driftObjectType"] == 0) | (p["driftSide"] == 0)].any()
    if not random:
        for f in ("k", "j", "i", "age", "health"):
            assert np.array_equal(dev[f], ref[f], equal_nan=True), f
        assert np.array_equal(dev["driftSide"], ref["driftSide"])
    else:
        # a particle that strands on a different substep (an ulp of the
        # device log/cos across the shore limit) stops flipping elsewhere:
        # the sides agree wherever the trajectories do
        ok = np.all(np.stack([close(dev[f], ref[f], 1e-9) for f in ("k", "j", "i", "age")]), axis=0)
        assert ok.mean() >= 0.97, float(ok.mean())
        assert np.array_equal(dev["driftSide"][ok], ref["driftSide"][ok])


def test_typed_particles_move_as_passive(ctx):
    """driftModel 0, no jibe: both paths move a typed particle exactly as the
    same particle made passive, and leave its type and side alone."""
    W = random_window(4)
    p = mixed_population(W, 3000, seed=5)
    passive = p.copy()
    passive["driftObjectType"], passive["driftSide"] = 0, 0
    cfg = wc.default_config(random=True, sigma=30.0, randomSeed=3)
    for path in ("parity", "fast"):
        a = run_device(ctx, W, cfg, p, 2, path=path)
        b = run_device(ctx, W, cfg, passive, 2, path=path)
        for f in ("id", "k", "j", "i", "health", "age", "time"):
            assert np.array_equal(a[f], b[f], equal_nan=f != "id"), (path, f)
        assert np.array_equal(a["driftObjectType"], p["driftObjectType"])
        assert np.array_equal(a["driftSide"], p["driftSide"])


@pytest.mark.parametrize("store", ["f32", "bf16"])
def test_fast_path_jibe_matches_parity(ctx, store):
    """The mixed-precision path flips sides with the reference's keyed draw:
    wherever a particle took as many live substeps as on the FP64 path (its
    age agrees), its side agrees bit for bit; its type never changes."""
    W = random_window(4)
    p = mixed_population(W, 3000, seed=6)
    # constraint closures and no survival threshold: a particle ends only on
    # an invalid cell, so most take every substep on both paths
    cfg = wc.default_config(random=True, sigma=30.0, survprob=0.0, tau0=20000.0, randomSeed=12, upperClosure=1,
                            lowerClosure=1, horizontalClosure=1, leewayJibeProbabilityHourly=0.5)
    ref = run_device(ctx, W, cfg, p, 3, path="parity")
    dev = run_device(ctx, W, cfg, p, 3, path="fast", store=store)
    assert np.array_equal(dev["driftObjectType"], p["driftObjectType"])
    same = dev["age"] == ref["age"]
    print(f"{store}: {same.mean():.4f} with equal ages", flush=True)
    assert same.mean() > 0.8
    assert np.array_equal(dev["driftSide"][same], ref["driftSide"][same])
    flipped = dev["driftSide"] != p["driftSide"]
    assert flipped.any() and not flipped[(p["driftObjectType"] == 0) | (p["driftSide"] == 0)].any()


def test_jibe_dropin(ctx):
    """wc_hip_move_particles (particle_data AoS, device pointers) with typed
    particles: the side outcomes of tests/ParticleMoveTest.cpp:192-215 at
    probability 1 and keyed flips at probability 0.4, against the oracle."""
    from hip_rt import DeviceBuffer
    host = dict(ocean=np.array([0.0, 65.0]), lon=np.array([0, DEG, 0, DEG]), lat=np.array([0, 0, DEG, DEG]),
                depth=np.array([.5, .5, .5]), h=np.full(4, 100.0), mask=np.ones(4),
                zeta=np.zeros(8, np.float32), u=np.zeros(16, np.float32), v=np.zeros(16, np.float32),
                w=np.zeros(24, np.float32), akt=np.zeros(24, np.float32))
    d = {k: DeviceBuffer(v) for k, v in host.items()}
    tw = tiny_window()
    L = wc.load_library()
    try:
        for prob, direction in ((1.0, 1), (1.0, -1), (0.4, 1), (0.4, -1)):
            cfg = move_cfg(leewayJibeProbabilityHourly=prob, trackingDirection=direction, randomSeed=11)
            t = 0 if direction > 0 else 1
            p = wc.make_particles(256, k=-.5, j=.25, i=.25, time=0.0 if direction > 0 else 65.0)
            p["id"] = np.arange(256, dtype=np.uint64) + np.uint64(2**53)
            p["driftObjectType"] = PERSON_IN_WATER
            p["driftSide"] = np.where(np.arange(256) % 2, LEFT, RIGHT)
            ref = p.copy()
            po.move(tw, cfg_to_oracle(cfg), t, ref)
            cfg_dev = DeviceBuffer(np.frombuffer(bytes(cfg), np.uint8))
            p_dev = DeviceBuffer(p)
            rc = L.wc_hip_move_particles(cfg_dev.ptr, p_dev.ptr, t, 2, 3, 2, 2, 2, d["ocean"].ptr,
                                         d["mask"].ptr, d["lon"].ptr, d["lat"].ptr, d["depth"].ptr, d["h"].ptr,
                                         d["zeta"].ptr, d["u"].ptr, d["v"].ptr, d["w"].ptr, d["akt"].ptr,
                                         None, None, None, None, len(p), 1, 1, None)
            assert rc == 0
            out = p_dev.get(p)
            cfg_dev.free()
            p_dev.free()
            assert np.array_equal(out["driftSide"], ref["driftSide"]), (prob, direction)
            assert np.array_equal(out["driftObjectType"], ref["driftObjectType"])
            if prob == 1.0:                                 # 3 live substeps: every side flipped
                assert np.array_equal(out["driftSide"], -p["driftSide"])
            for f in ("k", "j", "i", "age"):
                assert np.array_equal(out[f], ref[f]), f
    finally:
        for b in d.values():
            b.free()


def test_remove_dead_keeps_drift_state(ctx):
    rng = np.random.default_rng(12)
    W = as_window(tiny_window())
    from runner import load_grid
    load_grid(ctx, W)
    for n in (257, 100003):
        p = wc.make_particles(n, k=-.5, j=.5, i=.5)
        p["health"] = np.where(rng.random(n) < 0.3, -1.0, rng.random(n) + 0.01)
        p["driftObjectType"] = rng.integers(0, 12, n)
        p["driftSide"] = rng.choice([-1, 0, 1], n)
        ctx.set_particles(p)
        assert ctx.remove_dead() == int((p["health"] > 0).sum())
        assert np.array_equal(ctx.get_particles(), p[p["health"] > 0])


def test_ensemble_rejects_typed_particles(ctx):
    W = as_window(tiny_window())
    from runner import load_grid
    load_grid(ctx, W)
    p = wc.make_particles(4, k=-.5, j=.5, i=.5)
    p["driftObjectType"][2] = PERSON_IN_WATER
    with pytest.raises(wc.WcError) as e:
        ctx.ens_set(p, np.zeros(4, np.uint8), [1.0])
    assert e.value.status == 3

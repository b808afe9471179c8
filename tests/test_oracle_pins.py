"""Pins of the CPU oracle (oracle/wc_oracle.c) against the reference.

Each test restates a reference test, a golden value, a closed form or an
invariant; the planted-mistake tests at the end show that each pin fails on a
plausible error built into oracle/_plant/ (WCO_PLANT=n).

Reference tests restated here:
  tests/NumericalHelpersTest.cpp (helpers, RNG golden values),
  tests/ParticleMoveTest.cpp:10-139, 239-299 (passive move cases),
  tests/ConcentrationTest.cpp, tests/SourceEmissionTest.cpp,
  tests/ApplicationEndToEndTest.cpp:363-393 (uniform current on the sphere).
"""
import math

import numpy as np
import pytest

import pyoracle as po
from scenarios import DEG, haversine_1deg, rect_window, tiny_window

L = po.lib()
REF = po.ref_lib()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (reference sources absent)")

NAN = float("nan")


# --------------------------------------------------------------------------
# keyed RNG (NumericalHelpers.hpp:84-120)
# --------------------------------------------------------------------------

def test_rng_golden_values():
    # tests/NumericalHelpersTest.cpp:88-89
    assert L.wco_uniform(5489, 7, 3600, 2, 0) == float.fromhex("0x1.bd90c6590112ep-1")
    assert L.wco_uniform(0, 0, 0, 0, 0) == float.fromhex("0x1.e2b96a69ad7f4p-2")


def test_rng_key_order_properties():
    # tests/NumericalHelpersTest.cpp:63-75
    r0 = L.wco_normal(5489, 7, 3600, 2, 0)
    assert r0 == L.wco_normal(5489, 7, 3600, 2, 0)
    assert r0 != L.wco_normal(5490, 7, 3600, 2, 0)
    assert r0 != L.wco_normal(5489, 7, 3600, 3, 0)
    assert r0 != L.wco_normal(5489, 7, 3600, 2, 1)
    U = L.wco_uniform
    assert U(5489, 3, 3600, 7, 0) != U(5489, 7, 3600, 3, 0)
    assert U(5489, 5, 3600, 5, 0) != U(5489, 9, 3600, 9, 0)
    assert U(5489, 42, 3600, 2, 0) != U(42, 5489, 3600, 2, 0)
    assert U(5489, 42, 3600, 2, 0) != U(5489, 42, 3600, 0, 2)
    assert U(5489, 42, 3600, 5489, 0) != U(5489, 43, 3600, 5489, 0)


def test_rng_no_duplicates_lattice():
    # tests/NumericalHelpersTest.cpp:79-83 (1000 particles x 120 substeps x 6 components)
    draws = set()
    U = L.wco_uniform
    for particle in range(1000):
        for substep in range(120):
            for comp in range(6):
                draws.add(U(5489, particle, 1675900800, substep, comp))
    assert len(draws) == 1000 * 120 * 6


def test_normal_statistics():
    # Box-Muller output: mean 0, variance 1 (closed form of the transform)
    z = np.array([L.wco_normal(5489, p, 0, s, 0) for p in range(200) for s in range(100)])
    assert abs(z.mean()) < 4 / math.sqrt(z.size)
    assert abs(z.var() - 1) < 4 * math.sqrt(2 / z.size)


@needs_ref
def test_rng_matches_reference_build():
    assert REF.ref_scheme() == b"ordered-splitmix64-v2"
    rng = np.random.default_rng(1)
    for _ in range(3000):
        seed, part, sub, comp = (int(x) for x in rng.integers(0, 2**63, 4, dtype=np.uint64))
        interval = int(rng.integers(-2**62, 2**62))
        assert L.wco_mix(seed) == REF.ref_mix(seed)
        assert L.wco_uniform(seed, part, interval, sub, comp) == REF.ref_uniform(seed, part, interval, sub, comp)
        assert L.wco_normal(seed, part, interval, sub % 7, comp % 3) == \
            REF.ref_normal(seed, part, interval, sub % 7, comp % 3)


# --------------------------------------------------------------------------
# numerical helpers (NumericalHelpers.hpp:45-74)
# --------------------------------------------------------------------------

def test_helpers_reference_assertions():
    # tests/NumericalHelpersTest.cpp:25-55
    assert math.isnan(L.wco_restart_elapsed(0, 900, 900))
    assert L.wco_restart_elapsed(0, 900, 450) == 450
    assert math.isnan(L.wco_restart_elapsed(900, 0, 0))
    assert L.wco_restart_elapsed(900, 0, 450) == 450
    assert L.wco_restart_elapsed(2700, 5400, 2700) == 0
    assert math.isnan(L.wco_restart_elapsed(5400, 2700, 2700))
    assert abs(L.wco_reflect_cell(1.2, 1, .7, 0) - 1.5) < 1e-12
    assert abs(L.wco_reflect_cell(1.8, 1, 2.3, 2) - 1.7) < 1e-12
    assert L.wco_reflect_cell(1.2, 1, 1.4, 1) == 1.4
    assert abs(L.wco_reflect_domain(-.2, 2) - .2) < 1e-12
    assert abs(L.wco_reflect_domain(2.3, 2) - 1.7) < 1e-12
    assert abs(L.wco_reflect_domain(6.3, 2) - 1.7) < 1e-12
    assert L.wco_reflect_domain(2, 2) < 2


@needs_ref
def test_helpers_match_reference_build():
    rng = np.random.default_rng(2)
    for _ in range(2000):
        s, e, c = rng.uniform(-1e4, 1e4, 3)
        a, b = L.wco_restart_elapsed(s, e, c), REF.ref_restart_elapsed(s, e, c)
        assert (math.isnan(a) and math.isnan(b)) or a == b
        old = rng.uniform(0, 50)
        cand = old + rng.uniform(-2, 2)
        oc, cc = int(math.floor(old)), int(math.floor(cand))
        assert L.wco_reflect_cell(old, oc, cand, cc) == REF.ref_reflect_cell(old, oc, cand, cc)
        x, m = rng.uniform(-300, 300), float(rng.integers(-1, 120))
        assert L.wco_reflect_domain(x, m) == REF.ref_reflect_domain(x, m)


# --------------------------------------------------------------------------
# Particle::move (Particle.cpp:71-824): tests/ParticleMoveTest.cpp restated
# --------------------------------------------------------------------------

def move_cfg(**kw):
    # tests/ParticleMoveTest.cpp:24-31 (config_data{} zero-initialised: integrator 0 = Euler branch)
    base = dict(random=False, deltat=3600, dti=30, survprob=0, tau0=86400, crid=1, sigma=3.46,
                shoreLimit=.25, upperClosure=po.CONSTRAINT, lowerClosure=po.CONSTRAINT,
                horizontalClosure=po.CONSTRAINT, trackingDirection=1, restartCheckpoint=NAN,
                integrator=0, randomSeed=0)
    base.update(kw)
    return po.default_config(**base)


def one(id_, k, j, i, time, health=1.0, age=0.0):
    p = po.make_particles(1, k=k, j=j, i=i, time=time, health=health, age=age)
    p["id"] = id_
    return p


def u_ramp(value=2.0):
    u = np.zeros((2, 2, 2, 2), np.float32)
    u[1] = value
    return u


def case_time_linear(plant=0):
    """ParticleMoveTest.cpp:33-38: u ramps 0 -> 2 over 65 s (shortened last step)."""
    W = tiny_window(u=u_ramp())
    p = one(1, -.5, .25, .25, 0)
    po.move(W, move_cfg(), 0, p, plant=plant)
    xd = haversine_1deg()
    return abs(p["age"][0] - 65) < 1e-12 and abs(p["i"][0] - (.25 + 65.0 / xd)) < 1e-10, p


def test_move_time_linear_exact():
    ok, _ = case_time_linear()
    assert ok


def case_late_emission(plant=0):
    """ParticleMoveTest.cpp:40-44: emitted at t=40 into the 0 -> 2 m/s ramp."""
    W = tiny_window(u=u_ramp())
    p = one(2, -.5, .25, .25, 40)
    po.move(W, move_cfg(), 0, p, plant=plant)
    xd = haversine_1deg()
    return abs(p["age"][0] - 25) < 1e-12 and \
        abs(p["i"][0] - (.25 + (65.0 * 65.0 - 40.0 * 40.0) / 65.0 / xd)) < 1e-10


def test_move_late_emission():
    assert case_late_emission()


def test_move_backward_return():
    # ParticleMoveTest.cpp:46-50
    W = tiny_window(u=u_ramp())
    _, full = case_time_linear()
    p = one(3, -.5, full["j"][0], full["i"][0], 65)
    po.move(W, move_cfg(trackingDirection=-1), 1, p)
    assert abs(p["age"][0] - 65) < 1e-12
    assert abs(p["i"][0] - .25) < 1e-10


def test_move_seeded_repeatability():
    # ParticleMoveTest.cpp:52-64
    W = tiny_window(u=u_ramp())
    a, b = one(4, -.5, .25, .25, 0), one(4, -.5, .25, .25, 0)
    po.move(W, move_cfg(random=True, randomSeed=5489), 0, a)
    po.move(W, move_cfg(random=True, randomSeed=5489), 0, b)
    assert a["i"][0] == b["i"][0] and a["j"][0] == b["j"][0] and a["k"][0] == b["k"][0]
    c = one(4, -.5, .25, .25, 0)
    po.move(W, move_cfg(random=True, randomSeed=5490), 0, c)
    assert a["i"][0] != c["i"][0] or a["j"][0] != c["j"][0] or a["k"][0] != c["k"][0]


def test_move_stochastic_restart_equivalence():
    # ParticleMoveTest.cpp:66-81: a checkpoint at a substep boundary reproduces the draws
    W = tiny_window()
    cfg = move_cfg(random=True, randomSeed=5489)
    full = one(8, -.5, .25, .25, 0)
    po.move(W, cfg, 0, full)
    early = tiny_window(times=(0.0, 30.0))
    r = one(8, -.5, .25, .25, 0)
    po.move(early, cfg, 0, r)
    po.move(W, move_cfg(random=True, randomSeed=5489, restartCheckpoint=30), 0, r)
    for f in ("i", "j", "k", "age"):
        assert r[f][0] == full[f][0]


def test_move_parallel_equals_serial():
    # ParticleMoveTest.cpp:84-100 (OpenMP dynamic schedule vs serial, bit-exact)
    W = tiny_window()
    cfg = move_cfg(random=True, randomSeed=5489)
    a = po.make_particles(640, k=-.5, j=.25, i=.25, first_id=100)
    b = a.copy()
    po.move(W, cfg, 0, a, nthreads=1)
    po.move(W, cfg, 0, b, nthreads=8)
    assert np.array_equal(a, b)


def test_partition_ranges():
    # NumericalHelpers.hpp:76-82: the ranges tile [0, total) in rank order,
    # rank 0 takes the remainder, the others total // workers each
    for total in (0, 1, 10, 11, 4_000_000, 4_000_003):
        for workers in (1, 2, 3, 7, 8):
            ranges = [po.partition(total, workers, r) for r in range(workers)]
            assert ranges[0][0] == 0
            for (o0, n0), (o1, _) in zip(ranges, ranges[1:]):
                assert o1 == o0 + n0
            assert sum(n for _, n in ranges) == total
            assert all(n == total // workers for _, n in ranges[1:])
    # MpiDecompositionTest.cpp:15-20 with 11 particles over 4 ranks
    assert [po.partition(11, 4, r) for r in range(4)] == [(0, 5), (5, 2), (7, 2), (9, 2)]


@needs_ref
def test_partition_matches_reference_build():
    if not hasattr(REF, "ref_partition_count"):
        pytest.skip("oracle/_ref built before the partition helpers were added")
    for total in (0, 1, 11, 4_000_003, 2 ** 40 + 5):
        for workers in (1, 2, 3, 8, 13):
            for r in range(workers):
                assert po.partition(total, workers, r) == (REF.ref_partition_offset(total, workers, r),
                                                           REF.ref_partition_count(total, workers, r))


def test_move_decomposed_equals_serial():
    # MpiDecompositionTest.cpp:15-80: 11 particles scattered over the ranks by
    # partitionCount / partitionOffset, each range moved on its own, gathered
    # in rank order: bit-identical to one serial move, for every world size
    W = tiny_window(times=(0.0, 60.0))
    cfg = move_cfg(random=True, randomSeed=5489, deltat=60, dti=30, sigma=.1)
    p = po.make_particles(11, k=-.5, j=.25, i=.25, first_id=9007199254740993)
    serial = p.copy()
    po.move(W, cfg, 0, serial, nthreads=1)
    assert not np.array_equal(serial["i"], p["i"])
    for workers in (1, 2, 3, 4, 11):
        parts = []
        for r in range(workers):
            o, n = po.partition(len(p), workers, r)
            local = p[o:o + n].copy()
            po.move(W, cfg, 0, local, nthreads=1)
            parts.append(local)
        assert np.array_equal(np.concatenate(parts), serial)


# --------------------------------------------------------------------------
# drift object type and side: carried by the move, flipped by the hourly jibe
# (Particle.cpp:805-813, DriftModel.hpp:110-160)
# --------------------------------------------------------------------------

PERSON_IN_WATER, RIGHT, LEFT = 1, 1, -1


def typed(id_, k, j, i, time, side=RIGHT, kind=PERSON_IN_WATER):
    p = one(id_, k, j, i, time)
    p["driftObjectType"], p["driftSide"] = kind, side
    return p


def test_jibe_reference_cases():
    # tests/ParticleMoveTest.cpp:192-215 with driftModel 0: the side outcome
    # does not depend on the drift model (the positions there need leeway)
    W = tiny_window()                                      # ocean time 0 .. 65: substeps of 30, 30, 5 s
    cfg = move_cfg(leewayJibeProbabilityHourly=1.0)
    full = typed(30, -.5, .25, .25, 0)
    po.move(W, cfg, 0, full)
    r = typed(30, -.5, .25, .25, 0)
    po.move(tiny_window(times=(0.0, 30.0)), cfg, 0, r)
    po.move(W, move_cfg(leewayJibeProbabilityHourly=1.0, restartCheckpoint=30), 0, r)
    assert full["driftSide"][0] == LEFT and r["driftSide"][0] == full["driftSide"][0]
    assert full["driftObjectType"][0] == PERSON_IN_WATER
    # backward from 65, and restarted at 35 after the late window 35 .. 65
    back = move_cfg(leewayJibeProbabilityHourly=1.0, trackingDirection=-1)
    bf = typed(31, -.5, .25, .25, 65)
    po.move(W, back, 1, bf)
    br = typed(31, -.5, .25, .25, 65)
    po.move(tiny_window(times=(35.0, 65.0)), back, 1, br)
    po.move(W, move_cfg(leewayJibeProbabilityHourly=1.0, trackingDirection=-1, restartCheckpoint=35), 1, br)
    assert bf["driftSide"][0] == br["driftSide"][0] == LEFT
    # passive particles and an undefined side never change; no jibe without a probability
    passive = one(32, -.5, .25, .25, 0)
    undefined = typed(33, -.5, .25, .25, 0, side=0)
    po.move(W, cfg, 0, passive)
    po.move(W, cfg, 0, undefined)
    assert passive["driftSide"][0] == 0 and undefined["driftSide"][0] == 0
    off = typed(34, -.5, .25, .25, 0)
    po.move(W, move_cfg(), 0, off)
    assert off["driftSide"][0] == RIGHT


def test_jibe_flip_fraction_closed_form():
    # per substep the side flips with q = 1 - (1-p)^(dt/3600); after n
    # independent substeps it is flipped with probability (1 - (1-2q)^n)/2
    p_hour, n_part = 0.5, 20000
    W = tiny_window(times=(0.0, 3600.0))
    cfg = move_cfg(leewayJibeProbabilityHourly=p_hour, randomSeed=77)
    ps = po.make_particles(n_part, k=-.5, j=.25, i=.25, first_id=1000)
    ps["driftObjectType"], ps["driftSide"] = PERSON_IN_WATER, RIGHT
    po.move(W, cfg, 0, ps, nthreads=8)
    q = 1 - (1 - p_hour) ** (30 / 3600)
    expect = (1 - (1 - 2 * q) ** 120) / 2
    got = float(np.mean(ps["driftSide"] == LEFT))
    assert abs(got - expect) < 5 * math.sqrt(expect * (1 - expect) / n_part), (got, expect)
    assert set(np.unique(ps["driftSide"])) <= {LEFT, RIGHT}


@needs_ref
def test_drift_helpers_match_reference_build():
    if not hasattr(REF, "ref_jibe_step_probability"):
        pytest.skip("oracle/_ref built before the drift helpers were added")
    rng = np.random.default_rng(4)
    for h in (-0.5, 0.0, 1e-9, 0.05, 0.5, 0.999, 1.0, 2.0):
        for dt in (-1.0, 0.0, 1e-3, 5.0, 30.0, 3600.0, 7200.5):
            assert L.wco_jibe_step_probability(h, dt) == REF.ref_jibe_step_probability(h, dt)
    for _ in range(2000):
        seed, pid = (int(x) for x in rng.integers(0, 2 ** 63, 2))
        prob = float(rng.uniform(-0.1, 1.1))
        assert L.wco_sample_drift_side(seed, pid, prob) == REF.ref_sample_drift_side(seed, pid, prob)
    # the flip inside the move uses the reference's jibe component and step
    # probability on the substep key: recompute the side from those
    W = tiny_window(times=(0.0, 3600.0))
    cfg = move_cfg(leewayJibeProbabilityHourly=0.3, randomSeed=9)
    ps = po.make_particles(300, k=-.5, j=.25, i=.25, first_id=50)
    ps["driftObjectType"], ps["driftSide"] = PERSON_IN_WATER, RIGHT
    po.move(W, cfg, 0, ps, nthreads=1)
    prob = REF.ref_jibe_step_probability(0.3, 30.0)
    comp = REF.ref_jibe_component()
    for r in ps:
        side = RIGHT
        for sub in range(120):
            if REF.ref_uniform(9, int(r["id"]), 0, sub, comp) < prob:
                side = -side
        assert r["driftSide"] == side


def test_move_restart_forward_backward():
    # ParticleMoveTest.cpp:102-129
    _, full = case_time_linear()
    early_u = np.zeros((2, 2, 2, 2), np.float32)
    early_u[1] = np.float32(80.0) / np.float32(65.0)
    early = tiny_window(times=(0.0, 40.0), u=early_u)
    p = one(5, -.5, .25, .25, 0)
    po.move(early, move_cfg(), 0, p)
    po.move(tiny_window(u=u_ramp()), move_cfg(restartCheckpoint=40), 0, p)
    assert abs(p["i"][0] - full["i"][0]) < 1e-8
    assert abs(p["age"][0] - full["age"][0]) < 1e-12

    late_u = np.zeros((2, 2, 2, 2), np.float32)
    late_u[0] = np.float32(80.0) / np.float32(65.0)
    late_u[1] = 2.0
    late = tiny_window(times=(40.0, 65.0), u=late_u)
    q = one(6, -.5, full["j"][0], full["i"][0], 65)
    po.move(late, move_cfg(trackingDirection=-1), 1, q)
    po.move(tiny_window(u=u_ramp()), move_cfg(trackingDirection=-1, restartCheckpoint=40), 1, q)
    assert abs(q["i"][0] - .25) < 1e-8
    assert abs(q["age"][0] - 65) < 1e-12


def case_vertical(plant=0, k0=-.5):
    """ParticleMoveTest.cpp:133-139: w=2 m/s on the surface W level, 0 below,
    for 1 s.  Closed form: k1 = k0 + 2 (1 - beta) * 1 s / (h * dsigma), with
    beta = ceil(k0) - k0 the weight of the lower W level (k0=-.5 gives -.48)."""
    w = np.zeros((2, 3, 2, 2), np.float32)
    w[:, 2] = 2.0                       # logical level 0 -> storage index 2
    W = tiny_window(times=(0.0, 1.0), w=w)
    p = one(7, k0, .25, .25, 0)
    po.move(W, move_cfg(), 0, p, plant=plant)
    beta = math.ceil(k0) - k0
    return abs(p["k"][0] - (k0 + 2.0 * (1 - beta) / 50.0)) < 1e-12


def test_move_vertical_w():
    assert case_vertical()
    assert case_vertical(k0=-.25)


def shear_window(times=(0.0, 3600.0)):
    # ParticleMoveTest.cpp:250-252: u(t,k,j,i=1)=10 -> u = b*i with b = 10 m/s per cell
    u = np.zeros((2, 2, 2, 2), np.float32)
    u[:, :, :, 1] = 10.0
    return tiny_window(times=times, u=u)


def shear_error(integrator, dti, plant=0):
    cfg = move_cfg(integrator=integrator, dti=dti)
    p = one(40, -.5, .25, .25, 0)
    po.move(shear_window(), cfg, 0, p, plant=plant)
    assert p["health"][0] > 0
    exact = .25 * math.exp(10.0 * 3600.0 / haversine_1deg())
    return abs(p["i"][0] - exact)


def convergence_ok(plant=0):
    e600, e300 = shear_error(po.EULER, 600, plant), shear_error(po.EULER, 300, plant)
    m600, m300 = shear_error(po.MIDPOINT, 600, plant), shear_error(po.MIDPOINT, 300, plant)
    return (1.8 < e600 / e300 < 2.2) and (3.6 < m600 / m300 < 4.4) and (m600 < e600 / 20)


def test_move_integrator_orders():
    # ParticleMoveTest.cpp:239-266
    assert convergence_ok()


def test_move_midpoint_backward_and_restart():
    # ParticleMoveTest.cpp:268-287
    m300 = shear_error(po.MIDPOINT, 300)
    fwd = one(41, -.5, .25, .25, 0)
    po.move(shear_window(), move_cfg(integrator=po.MIDPOINT, dti=300), 0, fwd)
    back = one(42, -.5, fwd["j"][0], fwd["i"][0], 3600)
    po.move(shear_window(), move_cfg(integrator=po.MIDPOINT, dti=300, trackingDirection=-1), 1, back)
    assert abs(back["i"][0] - .25) < 2 * m300
    assert abs(back["age"][0] - 3600) < 1e-12
    r = one(41, -.5, .25, .25, 0)
    po.move(shear_window((0.0, 1800.0)), move_cfg(integrator=po.MIDPOINT, dti=300), 0, r)
    po.move(shear_window(), move_cfg(integrator=po.MIDPOINT, dti=300, restartCheckpoint=1800), 0, r)
    assert abs(r["i"][0] - fwd["i"][0]) < 1e-12
    assert abs(r["age"][0] - fwd["age"][0]) < 1e-12


def test_move_midpoint_euler_fallback():
    # ParticleMoveTest.cpp:289-299
    W = shear_window((0.0, 600.0))
    e = one(43, -.5, .25, .99, 0)
    po.move(W, move_cfg(integrator=po.EULER, dti=600), 0, e)
    m = one(43, -.5, .25, .99, 0)
    po.move(W, move_cfg(integrator=po.MIDPOINT, dti=600), 0, m)
    for f in ("i", "j", "k", "health"):
        assert m[f][0] == e[f][0]


def test_move_decay_and_survival():
    # Particle.cpp:218-223, 797-804: health = exp(-age/tau0); below survprob -> -1
    W = tiny_window()
    p = one(9, -.5, .25, .25, 0)
    po.move(W, move_cfg(tau0=100.0), 0, p)
    assert p["health"][0] == math.exp(-65.0 / 100.0)
    q = one(10, -.5, .25, .25, 0, health=0.5)
    po.move(W, move_cfg(survprob=0.6), 0, q)
    assert q["health"][0] == -1 and q["age"][0] == 0


def test_move_closures():
    W = tiny_window()
    # pre-step upper closure (Particle.cpp:134-153)
    for mode, expect_k, expect_h in ((po.CONSTRAINT, 0.0, None), (po.REFLECTION, -0.3, None), (po.KILL, None, -1)):
        p = one(11, 0.3, .25, .25, 0)
        po.move(W, move_cfg(upperClosure=mode), 0, p)
        if expect_k is not None:
            assert p["k"][0] == expect_k
        if expect_h is not None:
            assert p["health"][0] == -1
    # lower closure: kLowerLimit = -s_w+2 = -1 (Particle.cpp:125, 156-176)
    for mode, expect_k in ((po.CONSTRAINT, -1.0), (po.REFLECTION, -0.8)):
        p = one(12, -1.2, .25, .25, 0)
        po.move(W, move_cfg(lowerClosure=mode), 0, p)
        assert abs(p["k"][0] - expect_k) < 1e-15
    # open boundary (validInterpolationCell fails): kill (Particle.cpp:241-248)
    p = one(13, -.5, 1.2, .25, 0)
    po.move(W, move_cfg(), 0, p)
    assert p["health"][0] == -1


def test_move_horizontal_reflection_off_domain():
    # Particle.cpp:777-791: candidate off the grid -> reflectDomain on each axis
    u = np.full((2, 2, 2, 2), -200.0, np.float32)           # strong westward current
    W = tiny_window(times=(0.0, 30.0), u=u)
    p = one(14, -.5, .25, .01, 0)
    po.move(W, move_cfg(horizontalClosure=po.REFLECTION, integrator=po.EULER), 0, p)
    xd = haversine_1deg()
    cand = .01 - 200.0 * 30.0 / xd
    assert p["i"][0] == L.wco_reflect_domain(cand, 1)
    assert p["health"][0] > 0


def test_move_uniform_current_sphere():
    """ApplicationEndToEndTest.cpp:363-393 analogue on the plain CPU configuration
    grid (120x100x10 at 200 m, u=0.2, v=0.1): rhumb-line endpoint within 1e-3 of
    the path, k preserved to 1e-12, forward/backward round trip < 1 mm."""
    W = rect_window(120, 100, 10, 200.0, 40.75, 13, 0.2, 0.1)
    cfg = move_cfg(integrator=po.MIDPOINT, survprob=1e-4, lowerClosure=po.KILL,
                   horizontalClosure=po.REFLECTION)
    p = po.make_particles(4, k=-0.2, j=59.5, i=49.5, time=0.0)
    start = p.copy()
    for t in range(12):
        po.move(W, cfg, t, p)
    R = 6371000.0
    lat0, lon0 = W.latRad[0, 0], W.lonRad[0, 0]
    dlat, dlon = W.latRad[1, 0] - lat0, W.lonRad[0, 1] - lon0
    T = 12 * 3600.0
    phi0 = lat0 + 59.5 * dlat
    phi1 = phi0 + 0.1 * T / R
    lam0 = lon0 + 49.5 * dlon
    lam1 = lam0 + (0.2 / 0.1) * math.log(math.tan(math.pi / 4 + phi1 / 2) / math.tan(math.pi / 4 + phi0 / 2))
    path = math.hypot(0.2, 0.1) * T
    for q in p:
        lam, phi = lon0 + q["i"] * dlon, lat0 + q["j"] * dlat
        err = math.hypot((lam - lam1) * R * math.cos(phi1), (phi - phi1) * R)
        assert err < 1e-3 * path
        assert abs(q["k"] - (-0.2)) < 1e-12
        assert abs(q["age"] - T) < 1e-3
    back = p.copy()
    back["time"] = T
    bcfg = move_cfg(integrator=po.MIDPOINT, survprob=1e-4, lowerClosure=po.KILL,
                    horizontalClosure=po.REFLECTION, trackingDirection=-1)
    for t in range(12, 0, -1):
        po.move(W, bcfg, t, back)
    for q, s in zip(back, start):
        dx = (q["i"] - s["i"]) * dlon * R * math.cos(phi0)
        dy = (q["j"] - s["j"]) * dlat * R
        assert math.hypot(dx, dy) < 1e-3


def diffusion_variance(plant=0, n=20000):
    """docs/model.md:204-210: R = sigma_k Z sqrt(dt/dti), sigma_k = sigma (1 - k/k_b);
    with zero current the per-axis displacement variance after m full steps is
    m sigma_k^2.  Returns (measured/expected for x, for y)."""
    W = tiny_window(times=(0.0, 3600.0))
    sigma = math.sqrt(2 * 1.0 * 30.0)            # Kh = 1 m^2/s  ->  sigma = sqrt(2 Kh dti)
    cfg = move_cfg(random=True, randomSeed=5489, sigma=sigma, integrator=po.MIDPOINT)
    p = po.make_particles(n, k=-0.5, j=0.5, i=0.5)
    po.move(W, cfg, 0, p, plant=plant)
    xd = haversine_1deg()
    dx = (p["i"] - 0.5) * xd
    dy = (p["j"] - 0.5) * W.cell_metric(0, 0)[1]
    expected = 120 * (sigma * (1 - (-0.5) / (-1.0))) ** 2
    return dx.var() / expected, dy.var() / expected, dx.mean(), dy.mean(), expected


def test_move_diffusion_variance():
    rx, ry, mx, my, expected = diffusion_variance()
    tol = 5 * math.sqrt(2 / 20000)               # 5 standard errors of a variance estimate
    assert abs(rx - 1) < tol and abs(ry - 1) < tol
    assert abs(mx) < 5 * math.sqrt(expected / 20000) and abs(my) < 5 * math.sqrt(expected / 20000)


# --------------------------------------------------------------------------
# Concentration::evaluate (Concentration.hpp:16-44)
# --------------------------------------------------------------------------

def test_concentration_reference_case():
    # tests/ConcentrationTest.cpp:5-28
    n = 100000
    p = po.make_particles(n + 2, k=-1, j=1, i=1, time=0)
    p["j"][n] = 4                       # out of domain
    p["health"][n + 1] = 0              # dead (health 0 is not alive)
    c = po.grid_count(p, 3, 3, 3)
    assert c[-1 + 2, 1, 1] == n
    assert c.sum() == n


def rounding_case(plant=0):
    """(int)round is half away from zero (Concentration.hpp:31-33)."""
    ks = [-0.5, -1.5, -0.49, 0.0, -2.5, 0.4]
    js = [0.5, 1.5, 2.49, 0.0, 1.0, 1.0]
    is_ = [1.5, 0.49, 2.5, 2.0, 1.0, 1.0]
    p = po.make_particles(len(ks))
    p["k"], p["j"], p["i"] = ks, js, is_
    c = po.grid_count(p, 3, 4, 4, plant=plant)
    # expected cells: round half away from zero; (-2.5 -> -3) is below the grid
    want = {(-1, 1, 2): 1, (-2, 2, 0): 1, (0, 2, 3): 1, (0, 0, 2): 1, (0, 1, 1): 1}
    got = {(k - 2, j, i): int(c[k, j, i]) for k, j, i in zip(*np.nonzero(c))}
    return got == want


def test_concentration_rounding_and_bounds():
    assert rounding_case()


# --------------------------------------------------------------------------
# sourceEmissionTimes (SourceEmission.hpp:38-79): tests/SourceEmissionTest.cpp
# --------------------------------------------------------------------------

def test_emission_reference_cases():
    E = po.emission_times
    hourly = E(0, 1000, 0, -1, 0, 3600)
    assert len(hourly) == 1000 and abs(hourly[0] - 1.8) < 1e-12 and abs(hourly[-1] - 3598.2) < 1e-9
    gap = E(0, 1000, 0, -1, 3600, 10800)
    assert len(gap) == 2000 and abs(gap[0] - 3601.8) < 1e-9 and abs(gap[-1] - 10798.2) < 1e-9
    a, b = E(0, 7, 0, -1, 0, 450), E(0, 7, 0, -1, 450, 3600)
    assert len(a) + len(b) == 7
    assert np.array_equal(np.concatenate([a, b]), E(0, 7, 0, -1, 0, 3600))
    bounded = E(0, 10, 900, 2700, 0, 3600)
    assert len(bounded) == 5 and all(900 <= t < 2700 for t in bounded)
    legacy = E(1, 1000, 0, -1, 0, 7200)
    assert len(legacy) == 1000 and (legacy == 0).all()
    assert len(E(1, 4, 3600, 3600, 3600, 7200)) == 4
    assert len(E(1, 4, 3600, 3600, 7200, 10800)) == 0
    assert len(E(1, 4, 3600, 3600, 0, 3600)) == 0
    assert len(E(1, 4, 0, 7200, 7200, 10800)) == 4
    assert len(E(1, 4, 0, 1800, 3600, 7200)) == 0
    assert len(E(1, 4, 3600.00001, 3600.00001, 3600, 7200)) == 4
    assert len(E(1, 4, 3599.99999, 3599.99999, 3600, 7200)) == 4
    pulse = E(2, 25, 1800, -1, 0, 3600)
    assert len(pulse) == 25 and (pulse == 1800).all()
    assert len(E(2, 25, 1800, -1, 1800, 3600)) == 25
    assert len(E(2, 25, 1800, -1, 0, 1800)) == 0
    assert E(1, 1.5, 0, -1, 0, 3600) == -1


@needs_ref
def test_emission_matches_reference_build():
    rng = np.random.default_rng(3)
    for _ in range(300):
        mode = int(rng.integers(0, 3))
        value = float(rng.integers(0, 50)) if mode else float(rng.uniform(0, 3000))
        s0 = float(rng.uniform(0, 20000))
        s1 = -1.0 if rng.random() < .3 else s0 + float(rng.uniform(0, 20000))
        i0 = float(rng.integers(0, 10)) * 3600.0
        a = po.emission_times(mode, value, s0, s1, i0, i0 + 3600.0)
        b = po.emission_times(mode, value, s0, s1, i0, i0 + 3600.0, use_ref=True)
        if isinstance(a, int) or isinstance(b, int):
            assert a == b
        else:
            assert np.array_equal(a, b)


# --------------------------------------------------------------------------
# planted mistakes: every pin above must catch the mistake it targets
# --------------------------------------------------------------------------

def test_planted_time_weight_swap_is_caught():
    # a full interval integrates a linear ramp exactly either way round; the
    # late-emission case integrates only its second part and tells them apart
    assert not case_late_emission(plant=1)


def test_planted_wrong_rho_level_is_caught():
    # u differs between the two rho levels: the level used for u must be ceil(k)
    u = np.zeros((2, 2, 2, 2), np.float32)
    u[:, 1] = 1.0                       # logical level 0 only
    W = tiny_window(u=u)
    res = []
    for plant in (0, 2):
        p = one(1, -.5, .25, .25, 0)
        po.move(W, move_cfg(), 0, p, plant=plant)
        res.append(abs(p["i"][0] - (.25 + 65.0 / haversine_1deg())) < 1e-10)
    assert res == [True, False]


def test_planted_w_weight_swap_is_caught():
    # k=-.5 weighs both W levels equally; k=-.25 does not
    assert case_vertical(0, k0=-.25) and not case_vertical(plant=3, k0=-.25)


def test_planted_euler_time_level_is_caught():
    assert case_time_linear(0)[0] and not case_time_linear(plant=4)[0]


def test_planted_sigma_depth_is_caught():
    rx, ry, *_ = diffusion_variance(plant=5, n=4000)
    assert abs(rx - 1) > 0.5 and abs(ry - 1) > 0.5


def test_planted_midpoint_metric_is_caught():
    """The midpoint second stage samples a different cell with a different
    metric; Particle.cpp:627-629 rescales its leap by D(X)/D(X*).  Closed form
    with u=0 and a uniform v: the index advances by exactly v dt / Dy(midpoint
    cell), so dropping the rescaling (dividing by the start cell's Dy) fails."""
    eta, xi = 3, 2
    jj, ii = np.meshgrid(np.arange(eta), np.arange(xi), indexing="ij")
    lat = np.array([[0.0, 0.0], [1.0, 1.0], [3.0, 3.0]]) * DEG   # row 0 is 1 deg tall, row 1 is 2 deg
    lon = ii * DEG
    from pyoracle import Window
    v = np.full((2, 2, eta, xi), 400.0, np.float32)
    W = Window(np.array([0.0, 30.0]), lon, lat, np.full((eta, xi), 100.0), np.array([.5, .5]),
               np.zeros((2, eta, xi), np.float32), np.zeros((2, 2, eta, xi), np.float32), v,
               np.zeros((2, 3, eta, xi), np.float32), np.zeros((2, 3, eta, xi), np.float32))
    dy_mid = W.cell_metric(1, 0)[1]
    res = []
    for plant in (0, 6):
        p = one(5, -.5, 0.99, .25, 0)
        po.move(W, move_cfg(integrator=po.MIDPOINT), 0, p, plant=plant)
        res.append(abs(p["j"][0] - (0.99 + 400.0 * 30.0 / dy_mid)) < 1e-12)
    assert res == [True, False]


def test_planted_rounding_is_caught():
    assert rounding_case(0) and not rounding_case(plant=7)

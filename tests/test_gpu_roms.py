"""GPU parity of the ROMS normalisation (wc_set_roms_grid +
wc_upload_roms_record, ROMSAdapter.cpp:56-300, 309-368, 448-466) against the
oracle's restatement, through the C ABI.

Bar: bit-identical records (the kernel keeps the reference's float sums,
float mean, double rotation and float rounding; the angle's cos/sin are taken
on the host with the same libm as the oracle), the rejections of
tests/ROMSAdapterTest.cpp, and an end-to-end move over a device-normalised
window equal to the oracle's move over the oracle-normalised window.
"""
import numpy as np
import pytest

import pyoracle as po
import wacommplusplus_amd as wc
from roms_cases import NC_FILL_FLOAT, fixture, oracle_grid, oracle_record, random_roms
from runner import cfg_to_oracle
from wacommplusplus_amd.synthetic import ForcingWindow

pytestmark = pytest.mark.gpu

PI = np.pi


@pytest.fixture(scope="module")
def ctx():
    c = wc.Context(0)
    yield c
    c.close()


def load(ctx, c, earth_relative=False, missing_u=(), missing_v=(), missing_angle=(), n_slots=None):
    T = c["u"].shape[0]
    ctx.set_roms_grid(c["s_w"], c["lon"], c["lat"], c["h"], c["mask_rho"], c["mask_u"], c["mask_v"],
                      angle=c["angle"], earth_relative=earth_relative, missing_u=missing_u, missing_v=missing_v,
                      missing_angle=missing_angle, n_slots=n_slots or max(2, T))


def upload(ctx, c, t, slot=None):
    ctx.upload_roms_record(t if slot is None else slot, c["zeta"][t], c["u"][t], c["v"][t], c["w"][t], c["akt"][t])


def assert_records_equal(ctx, c, earth_relative=False, **kw):
    load(ctx, c, earth_relative=earth_relative, **kw)
    g = oracle_grid(po, c, earth_relative=earth_relative, **kw)
    for t in range(c["u"].shape[0]):
        upload(ctx, c, t)
        got = ctx.get_record(t)
        ref = oracle_record(po, g, c, t)
        assert ref[0] == 0
        for name, a, b in zip(("zeta", "u", "v", "w", "akt"), got, ref[1:]):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (name, t)
    return got


def test_fixture_matches_reference_expectations(ctx):
    # tests/ROMSAdapterTest.cpp:67-74
    c = fixture()
    zeta, u, v, w, akt = assert_records_equal(ctx, c)
    assert np.all(u[:, :, 0] == 2) and np.all(u[:, :, 1] == 3) and np.all(u[:, :, 2] == 4) and np.all(v == 6)


@pytest.mark.parametrize("theta", [0.0, 0.5 * PI, -0.5 * PI, PI, 0.25 * PI])
def test_fixture_rotation(ctx, theta):
    # tests/ROMSAdapterTest.cpp:82-93
    c = fixture()
    c["angle"] = np.full((2, 3), theta)
    zeta, u, v, w, akt = assert_records_equal(ctx, c)
    for i in range(3):
        assert np.all(np.abs(u[:, :, i] - ((2 + i) * np.cos(theta) - 6 * np.sin(theta))) < 1e-6)
        assert np.all(np.abs(v[:, :, i] - ((2 + i) * np.sin(theta) + 6 * np.cos(theta))) < 1e-6)


def test_fixture_varied_angles_with_land(ctx):
    # tests/ROMSAdapterTest.cpp:104-115
    c = fixture()
    c["angle"] = np.array([0, 0.5 * PI, -0.5 * PI, PI, 0.25 * PI, np.nan]).reshape(2, 3)
    c["mask_rho"] = np.array([1, 1, 1, 1, 1, 0.0]).reshape(2, 3)
    zeta, u, v, w, akt = assert_records_equal(ctx, c)
    assert abs(u[0, 0, 1] + 6) < 1e-6 and abs(v[0, 0, 1] - 3) < 1e-6
    assert u[0, 1, 2] == 0 and v[0, 1, 2] == 0


def test_fixture_earth_relative(ctx):
    # tests/ROMSAdapterTest.cpp:147-160
    c = fixture()
    c["angle"] = np.full((2, 3), PI)
    zeta, u, v, w, akt = assert_records_equal(ctx, c, earth_relative=True)
    assert u[0, 0, 1] == 3 and v[0, 0, 1] == 6


@pytest.mark.parametrize("seed,earth", [(7, False), (8, False), (9, True)])
def test_random_grid_bit_identical(ctx, seed, earth):
    c = random_roms(seed=seed, T=3, s_rho=5, eta=48, xi=64)
    assert_records_equal(ctx, c, earth_relative=earth)


def test_large_ragged_grid_bit_identical(ctx):
    """A plane that is not a multiple of the block size, 300x257x7."""
    c = random_roms(seed=11, T=1, s_rho=7, eta=300, xi=257)
    assert_records_equal(ctx, c)


def expect_rejected(fn, text):
    with pytest.raises(wc.WcError) as e:
        fn()
    assert e.value.status == wc.WC_ERR_INVALID and text in str(e.value), str(e.value)


def test_rejections(ctx):
    c = fixture()
    # angle: NaN, inf, NetCDF fill, the missing_value attribute at a wet point (:116-131)
    for bad in (np.nan, np.inf, 9.9692099683868690e+36, -9999.0):
        c["angle"] = np.array([0, 0, 0, 0, 0, bad]).reshape(2, 3)
        expect_rejected(lambda: load(ctx, c, missing_angle=(-9999.0,)), "angle")
    c["angle"][:] = 0.0
    # a mask value other than 0/1 (:147-148, 154-155)
    c["mask_u"][1, 1] = 0.5
    expect_rejected(lambda: load(ctx, c), "mask")
    # NaN u at a wet face (:180-187); the slot stays unusable afterwards
    c = fixture()
    c["u"][0, 0, 0, 0] = np.nan
    load(ctx, c)
    upload(ctx, c, 1)
    expect_rejected(lambda: upload(ctx, c, 0), "u is missing")
    cfg = wc.default_config(random=False)
    expect_rejected(lambda: ctx.move_interval(cfg, 0.0, 3600.0, 0, 1), "not uploaded")
    # v: NetCDF float fill and an attribute missing value at wet faces
    c = fixture()
    c["v"][0, 1, 0, 2] = NC_FILL_FLOAT
    load(ctx, c)
    expect_rejected(lambda: upload(ctx, c, 0), "v is missing")
    c = fixture()
    c["v"][0, 0, 0, 1] = -32767.0
    load(ctx, c, missing_v=(-32767.0,))
    expect_rejected(lambda: upload(ctx, c, 0), "v is missing")
    # overflow of the float mean -> rotated velocity outside float range (:290-298)
    c = fixture()
    c["u"][0] = np.float32(3.4e38)
    load(ctx, c)
    expect_rejected(lambda: upload(ctx, c, 0), "float range")
    # the same bad values on land faces are accepted
    c = fixture()
    c["mask_u"][0, 0] = 0.0
    c["u"][0, :, 0, 0] = np.nan
    assert_records_equal(ctx, c)


def roms_window(c, g):
    """The oracle-normalised window as a ForcingWindow (static fields of
    ROMSAdapter.cpp:232-275 via wco_roms_static)."""
    lr, la, h, di = po.roms_static(c["s_w"], c["lon"], c["lat"], c["h"], c["mask_rho"])
    recs = {}

    def rec(t):
        if t not in recs:
            rc, Z, U, V, W, A = oracle_record(po, g, c, t)
            assert rc == 0
            recs[t] = (Z, U, V, W, A)
        return recs[t]
    s_rho = len(c["s_w"]) - 1
    eta, xi = c["mask_rho"].shape
    return ForcingWindow(s_rho, eta, xi, lr, la, h, c["mask_rho"], di, c["times"], rec)


@pytest.mark.parametrize("random", [False, True])
def test_move_over_roms_window(ctx, random):
    """Particles moved over a device-normalised ROMS window equal the
    oracle's move over the oracle-normalised one: bit-identical without
    diffusion; with it, >= 99.5 % of particles within 1e-9 cells (an ulp of
    the device's FP64 log/cos can flip a reflection decision)."""
    c = random_roms(seed=21, T=3, s_rho=5, eta=48, xi=64, land=0.15)
    g = oracle_grid(po, c)
    W = roms_window(c, g)
    rng = np.random.default_rng(4)
    p = wc.make_particles(3000, k=-0.5)
    p["j"] = rng.uniform(2.0, 45.0, len(p))
    p["i"] = rng.uniform(0.2, 1.0, len(p))
    cfg = wc.default_config(random=random, sigma=wc.sigma_from_kh(1.0, 30.0))
    ref = p.copy()
    ocfg = cfg_to_oracle(cfg)
    for t in range(2):
        r0, r1 = W.record(t), W.record(t + 1)
        win = po.Window(W.times[t:t + 2], W.lonRad, W.latRad, W.h, W.depthIntervals,
                        *[np.stack([r0[q], r1[q]]) for q in range(5)])
        po.move(win, ocfg, 0, ref)
    load(ctx, c)
    for t in range(3):
        upload(ctx, c, t)
    ctx.set_particles(p)
    for t in range(2):
        ctx.move_interval(cfg, float(W.times[t]), float(W.times[t + 1]), t, t + 1)
    dev = ctx.get_particles()
    ok = (dev["health"] > 0) == (ref["health"] > 0)
    for f in ("k", "j", "i"):
        d = np.abs(dev[f] - ref[f])
        d[np.isnan(dev[f]) & np.isnan(ref[f])] = 0.0
        ok &= d <= (1e-9 if random else 0.0)
    print(f"diffusion={'on' if random else 'off'}: {ok.mean():.5f} of particles match", flush=True)
    assert ok.mean() >= (0.995 if random else 1.0)


# --------------------------------------------------------------------------
# appendBoundaryRecord checks (OceanModelAdapter.cpp:199-233)
# --------------------------------------------------------------------------

def first_conflict(a, b):
    """Brute force of the reference's loop order (OceanModelAdapter.cpp:
    218-231): points in (j, i) order; per point zeta, then u/v over the
    levels, then w/AKt; the first conflict under float !=."""
    za, ua, va, wa, ka = a
    zb, ub, vb, wb, kb = b
    cls = [za != zb, np.any((ua != ub) | (va != vb), axis=0), np.any((wa != wb) | (ka != kb), axis=0)]
    for p in range(za.size):
        for q in range(3):
            if cls[q].flat[p]:
                return q + 1, p
    return 0, -1


def test_compare_records_duplicate_boundary(ctx):
    c = random_roms(seed=31, T=2, s_rho=4, eta=24, xi=32)
    load(ctx, c, n_slots=3)
    upload(ctx, c, 0, slot=0)
    upload(ctx, c, 0, slot=1)
    assert ctx.compare_records(0, 1) == (0, -1)
    base = ctx.get_record(0)
    m, mu, mv = c["mask_rho"], c["mask_u"], c["mask_v"]
    wet = np.argwhere(m > 0)
    # a wet point whose u face (j, i) and v face (j, i) are wet too
    cand = [(int(a), int(b)) for a, b in wet if b < mu.shape[1] and a < mv.shape[0] and mu[a, b] and mv[a, b]]
    j, i = cand[len(cand) // 2]
    # tests/NativeAdapterTest.cpp:77-90: each field + 1 at one point is a conflict of its class
    for name, cls in (("zeta", 1), ("u", 2), ("v", 2), ("w", 3), ("akt", 3)):
        d = {q: c[q].copy() for q in ("zeta", "u", "v", "w", "akt")}
        if name == "zeta":
            d["zeta"][0, j, i] += 1.0
        elif name == "u":
            d["u"][0, 1, j, i] += 1.0
        elif name == "v":
            d["v"][0, 2, j, i] += 1.0
        else:
            d[name][0, 3, j, i] += 1.0
        ctx.upload_roms_record(2, d["zeta"][0], d["u"][0], d["v"][0], d["w"][0], d["akt"][0])
        got = ctx.compare_records(0, 2)
        assert got == first_conflict(base, ctx.get_record(2)), (name, got)
        assert got[0] == cls, (name, got)
    # a later-class conflict at an earlier point wins over an earlier class at a later point
    d = {q: c[q].copy() for q in ("zeta", "u", "v", "w", "akt")}
    d["zeta"][0, -1, -1] = 5.0
    d["w"][0, 0, wet[0][0], wet[0][1]] += 1.0
    ctx.upload_roms_record(2, d["zeta"][0], d["u"][0], d["v"][0], d["w"][0], d["akt"][0])
    got = ctx.compare_records(0, 2)
    assert got == first_conflict(base, ctx.get_record(2)) and got[0] == 3
    # different records: the first point in loop order
    upload(ctx, c, 1, slot=2)
    assert ctx.compare_records(0, 2) == first_conflict(base, ctx.get_record(2))
    # NaN never equals itself
    d = {q: c[q].copy() for q in ("zeta", "u", "v", "w", "akt")}
    d["akt"][0, :, j, i] = np.nan
    ctx.upload_roms_record(1, d["zeta"][0], d["u"][0], d["v"][0], d["w"][0], d["akt"][0])
    ctx.upload_roms_record(2, d["zeta"][0], d["u"][0], d["v"][0], d["w"][0], d["akt"][0])
    assert ctx.compare_records(1, 2) == (3, j * c["mask_rho"].shape[1] + i)


def test_check_roms_grid(ctx):
    c = random_roms(seed=33, T=1, s_rho=4, eta=24, xi=32)
    s_rho_values = np.linspace(-0.875, -0.125, 4)
    ctx.set_roms_grid(c["s_w"], c["lon"], c["lat"], c["h"], c["mask_rho"], c["mask_u"], c["mask_v"],
                      angle=c["angle"], s_rho_values=s_rho_values)

    def check(**over):
        a = dict(s_w=c["s_w"], lon_rho=c["lon"], lat_rho=c["lat"], h=c["h"], mask_rho=c["mask_rho"],
                 mask_u=c["mask_u"], mask_v=c["mask_v"], s_rho_values=s_rho_values)
        a.update(over)
        ctx.check_roms_grid(**a)

    check()
    check(s_rho_values=None)
    land = np.argwhere(c["mask_rho"] == 0)[0]
    wet = np.argwhere(c["mask_rho"] > 0)[0]
    ok = [dict(s_w=c["s_w"] + 1e-13), dict(s_rho_values=s_rho_values + 1e-13), dict(lon_rho=c["lon"] + 1e-11),
          dict(lat_rho=c["lat"] - 1e-11), dict(h=c["h"] + 1e-9)]
    h_land = c["h"].copy()
    h_land[tuple(land)] += 5.0                       # land h is zeroed on both sides
    ok.append(dict(h=h_land))
    for o in ok:
        check(**o)
    h_wet = c["h"].copy()
    h_wet[tuple(wet)] += 1e-7
    mask = c["mask_rho"].copy()
    mask[tuple(land)] = 1.0
    bad = [(dict(s_w=c["s_w"] + 1e-11), "vertical coordinates"),
           (dict(s_rho_values=s_rho_values - 1e-11), "vertical coordinates"),
           (dict(lon_rho=c["lon"] + 1e-9), "horizontal grids"), (dict(lat_rho=c["lat"] + 1e-9), "horizontal grids"),
           (dict(h=h_wet), "horizontal grids"), (dict(mask_rho=mask), "horizontal grids"),
           (dict(lon_rho=c["lon"][:-1], lat_rho=c["lat"][:-1], h=c["h"][:-1], mask_rho=c["mask_rho"][:-1],
                 mask_v=c["mask_v"][:-1]), "grid dimensions")]
    for o, text in bad:
        expect_rejected(lambda: check(**o), text)


def test_run_window_rejects_non_chronological_time(ctx):
    # ROMSAdapter.cpp:197-201 / OceanModelAdapter.cpp:264-265
    c = random_roms(seed=35, T=3, s_rho=2, eta=8, xi=10)
    g = oracle_grid(po, c)
    W = roms_window(c, g)
    ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask)
    zeta, u, v, w, akt = W.stacked()
    cfg = wc.default_config(random=False)
    for times in ([0.0, 3600.0, 3600.0], [0.0, 7200.0, 3600.0], [0.0, np.nan, 7200.0]):
        expect_rejected(lambda: ctx.run_window(cfg, np.array(times), zeta, u, v, w, akt), "chronological")

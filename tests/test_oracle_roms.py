"""Pins of the oracle's ROMS normalisation (wco_roms_normalize,
wco_roms_static; ROMSAdapter.cpp:29-49, 56-300, 309-368, 448-466).

Restated from tests/ROMSAdapterTest.cpp: the u/v averaging of its fixture
(:67-74), the rotation by five uniform angles (:82-93), per-point angles with
a land point (:104-115), the rejections of a NaN / infinite / fill / missing
angle at a wet point (:116-131) and of a NaN u at a wet face (:180-187), and
earth-relative velocities that are not rotated (:147-160).  Closed forms
beyond those: a linear u field averages to face midpoints, rotation keeps the
speed and adds the angle, a dry face drops out of the mean.
"""
import math

import numpy as np

import pyoracle as po
from roms_cases import NC_FILL_DOUBLE, NC_FILL_FLOAT, fixture, oracle_grid, oracle_record, random_roms

PI = math.acos(-1.0)


def test_fixture_average():
    # tests/ROMSAdapterTest.cpp:67-74
    c = fixture()
    g = oracle_grid(po, c)
    for t in range(2):
        rc, Z, U, V, W, A = oracle_record(po, g, c, t)
        assert rc == 0
        assert np.all(U[:, :, 0] == 2) and np.all(U[:, :, 1] == 3) and np.all(U[:, :, 2] == 4)
        assert np.all(V == 6)


def test_fixture_rotation():
    # tests/ROMSAdapterTest.cpp:82-93
    c = fixture()
    for theta in (0.0, 0.5 * PI, -0.5 * PI, PI, 0.25 * PI):
        c["angle"] = np.full((2, 3), theta)
        g = oracle_grid(po, c)
        rc, Z, U, V, W, A = oracle_record(po, g, c, 0)
        assert rc == 0
        for i in range(3):
            u, v = 2 + i, 6
            assert np.all(np.abs(U[:, :, i] - (u * math.cos(theta) - v * math.sin(theta))) < 1e-6)
            assert np.all(np.abs(V[:, :, i] - (u * math.sin(theta) + v * math.cos(theta))) < 1e-6)
            assert np.all(np.abs(np.hypot(U[:, :, i], V[:, :, i]) - math.hypot(u, v)) < 1e-6)


def test_fixture_varied_angles_with_land():
    # tests/ROMSAdapterTest.cpp:104-115: storage level 0 is k=-1
    c = fixture()
    c["angle"] = np.array([0, 0.5 * PI, -0.5 * PI, PI, 0.25 * PI, math.nan]).reshape(2, 3)
    c["mask_rho"] = np.array([1, 1, 1, 1, 1, 0.0]).reshape(2, 3)
    rc, Z, U, V, W, A = oracle_record(po, oracle_grid(po, c), c, 0)
    assert rc == 0
    assert abs(U[0, 0, 1] + 6) < 1e-6 and abs(V[0, 0, 1] - 3) < 1e-6
    assert U[0, 1, 2] == 0 and V[0, 1, 2] == 0


def test_fixture_rejections():
    c = fixture()
    # NaN angle at a wet point (:116-120)
    c["angle"] = np.array([0, 0.5 * PI, -0.5 * PI, PI, 0.25 * PI, math.nan]).reshape(2, 3)
    assert oracle_record(po, oracle_grid(po, c), c, 0)[0] == -2
    # infinite, NetCDF double fill, and the missing_value attribute (:121-131)
    for bad in (math.inf, NC_FILL_DOUBLE, -9999.0):
        c["angle"] = np.array([0, 0, 0, 0, 0, bad]).reshape(2, 3)
        assert oracle_record(po, oracle_grid(po, c, missing_angle=(-9999.0,)), c, 0)[0] == -2
    # the same values on a land point are accepted
    c["mask_rho"] = np.array([1, 1, 1, 1, 1, 0.0]).reshape(2, 3)
    assert oracle_record(po, oracle_grid(po, c, missing_angle=(-9999.0,)), c, 0)[0] == 0
    # NaN u at a wet face (:180-187)
    c = fixture()
    c["u"][0, 0, 0, 0] = np.nan
    assert oracle_record(po, oracle_grid(po, c), c, 0)[0] == -3
    # a v fill value at a wet face
    c = fixture()
    c["v"][0, 1, 0, 2] = NC_FILL_FLOAT
    assert oracle_record(po, oracle_grid(po, c), c, 0)[0] == -4
    # a mask value other than 0/1 (:147-148)
    c = fixture()
    c["mask_u"][1, 1] = 0.5
    assert oracle_record(po, oracle_grid(po, c), c, 0)[0] == -1
    # a rotated velocity beyond float range: FLT_MAX on both faces overflows the float sum
    c = fixture()
    c["u"][0, :, :, :] = np.float32(3.4e38)
    assert oracle_record(po, oracle_grid(po, c), c, 0)[0] == -5


def test_fixture_earth_relative_ignores_angle():
    # tests/ROMSAdapterTest.cpp:147-160
    c = fixture()
    c["angle"] = np.full((2, 3), PI)
    rc, Z, U, V, W, A = oracle_record(po, oracle_grid(po, c, earth_relative=True), c, 0)
    assert rc == 0 and U[0, 0, 1] == 3 and V[0, 0, 1] == 6


def test_linear_field_averages_to_face_midpoints():
    """u(i) = a + b*i on the u grid: interior rho points get the mean of
    faces i-1 and i, edge points the single face beside them; v likewise in j."""
    eta, xi, s_rho = 5, 7, 2
    a, b = 0.25, 0.125                         # all sums and halves exact in float
    c = dict(mask_rho=np.ones((eta, xi)), mask_u=np.ones((eta, xi - 1)), mask_v=np.ones((eta - 1, xi)),
             angle=np.zeros((eta, xi)), s_w=np.linspace(-1, 0, s_rho + 1))
    u = np.broadcast_to(a + b * np.arange(xi - 1), (s_rho, eta, xi - 1)).astype(np.float32)
    v = np.broadcast_to((a + b * np.arange(eta - 1))[:, None], (s_rho, eta - 1, xi)).astype(np.float32)
    z, w = np.zeros((eta, xi), np.float32), np.zeros((s_rho + 1, eta, xi), np.float32)
    rc, Z, U, V, W, A = po.roms_normalize(oracle_grid(po, c), z, u, v, w, w)
    assert rc == 0
    expect_u = np.r_[a, a + b * (np.arange(1, xi - 1) - 0.5), a + b * (xi - 2)]
    expect_v = np.r_[a, a + b * (np.arange(1, eta - 1) - 0.5), a + b * (eta - 2)]
    assert np.array_equal(U[0], np.broadcast_to(expect_u, (eta, xi)).astype(np.float32))
    assert np.array_equal(V[1], np.broadcast_to(expect_v[:, None], (eta, xi)).astype(np.float32))


def test_dry_face_drops_out():
    c = random_roms(seed=3, T=1, s_rho=2, eta=6, xi=8, land=0.0, flips=0)
    c["mask_u"][2, 3] = 0.0                    # face between rho (2,3) and (2,4)
    c["mask_v"][1, 5] = 0.0                    # face between rho (1,5) and (2,5)
    c["angle"][:] = 0.0
    rc, Z, U, V, W, A = oracle_record(po, oracle_grid(po, c), c, 0)
    assert rc == 0
    u, v = c["u"][0], c["v"][0]
    assert np.array_equal(U[:, 2, 3], u[:, 2, 2]) and np.array_equal(U[:, 2, 4], u[:, 2, 4])
    assert np.array_equal(V[:, 1, 5], v[:, 0, 5]) and np.array_equal(V[:, 2, 5], v[:, 2, 5])


def test_rotation_keeps_speed_and_adds_angle():
    c = random_roms(seed=5, T=1)
    g0 = oracle_grid(po, c, earth_relative=True)
    g1 = oracle_grid(po, c)
    _, _, U0, V0, _, _ = oracle_record(po, g0, c, 0)
    rc, Z, U1, V1, W, A = oracle_record(po, g1, c, 0)
    assert rc == 0
    wet = np.broadcast_to(c["mask_rho"] > 0, U0.shape) & (np.hypot(U0, V0) > 1e-3)
    speed0, speed1 = np.hypot(U0, V0)[wet], np.hypot(U1, V1)[wet]
    assert np.max(np.abs(speed1 - speed0) / speed0) < 1e-6
    dphi = np.angle(np.exp(1j * (np.arctan2(V1, U1) - np.arctan2(V0, U0) -
                                 np.broadcast_to(c["angle"], U0.shape))))[wet]
    assert np.max(np.abs(dphi)) < 1e-5


def test_land_masking_and_zeta_fill():
    c = random_roms(seed=9, T=1)
    rc, Z, U, V, W, A = oracle_record(po, oracle_grid(po, c), c, 0)
    assert rc == 0                              # fill values sit on land faces only
    land = c["mask_rho"] == 0
    assert np.all(U[:, land] == 0) and np.all(V[:, land] == 0)
    assert np.all(W[:, land] == 0) and np.all(A[:, land] == 0) and np.all(Z[land] == 0)
    assert np.array_equal(W[:, ~land], c["w"][0][:, ~land]) and np.array_equal(Z[~land], c["zeta"][0][~land])
    # a wet zeta at or above 1e36 is a fill value too (ROMSAdapter.cpp:258)
    c["zeta"][0, ~land] = np.where(np.arange((~land).sum()) % 2 == 0, np.float32(1e36), c["zeta"][0, ~land])
    Z2 = oracle_record(po, oracle_grid(po, c), c, 0)[1]
    assert np.all(Z2[~land][::2] == 0)


def test_static_fields():
    """ROMSAdapter.cpp:232-275: the fixture's depth intervals, radians with
    the reference's 0.0174533 factor, h zeroed on land."""
    c = fixture()
    c["mask_rho"][1, 2] = 0.0
    lr, la, h, di = po.roms_static(c["s_w"], c["lon"], c["lat"], c["h"], c["mask_rho"])
    assert np.array_equal(di, [0.5, 0.5])
    assert lr[0, 1] == 0.0174533 * 11 and la[1, 0] == 0.0174533 * 41
    assert h[1, 2] == 0 and h[0, 0] == 100

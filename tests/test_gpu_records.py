"""GPU parity of the particle history records (wc_particle_records:
Particles::netcdfRecords, Particles.cpp:115-147, with
OceanModelAdapter::kji2deplatlon, OceanModelAdapter.cpp:271-321) against the
oracle, through the C ABI, and the device -> restart file -> context round
trip (Particles::saveAsNetCDF / loadFromNetCDF), drift object type and side
included.

Bar: bit-identical.  The kept set and its order equal the oracle's, every
record's depth, lat and lon has the oracle's bits (FP64, no contraction, the
reference's term order), and id, k, j, i, health, age, time are the
particle's.  Inputs include positions on every side of the reference's range
checks, NaN and infinite coordinates, on a curvilinear ROMS grid and on the
full-size 1024x768x11 grid with 4M particles.
"""
import numpy as np
import pytest

import pyoracle as po
import wacommplusplus_amd as wc
from roms_cases import random_roms
from wacommplusplus_amd.synthetic import full_window

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = wc.Context(0)
    yield c
    c.close()


def positions(n, s_w, eta, xi, seed):
    """A mix of in-domain and out-of-domain positions with edge cases."""
    rng = np.random.default_rng(seed)
    p = wc.make_particles(n)
    p["k"] = rng.uniform(-(s_w + 0.5), 0.5, n)
    p["j"] = rng.uniform(-2.0, eta + 1.0, n)
    p["i"] = rng.uniform(-2.0, xi + 1.0, n)
    p["health"] = rng.uniform(0, 1, n)
    p["age"] = rng.uniform(0, 1e5, n)
    p["time"] = rng.uniform(0, 1e5, n)
    p["id"] = rng.permutation(n).astype(np.uint64) + np.uint64(2 ** 53)
    p["driftObjectType"] = np.where(rng.random(n) < 0.5, 0, rng.integers(1, 12, n))
    p["driftSide"] = rng.choice([-1, 0, 1], n)
    edge = [(0.0, 0.0, 0.0), (-0.0, -0.0, -0.0), (-(s_w - 2.0), 1.0, 1.0), (-(s_w - 1.0), 1.0, 1.0),
            (-0.5, eta - 1.0, 2.0), (-0.5, 2.0, xi - 1.0), (-0.5, eta - 1.0 - 1e-12, xi - 1.0 - 1e-12),
            (np.nan, 1.0, 1.0), (-0.5, np.nan, 1.0), (-0.5, 1.0, np.inf), (-np.inf, 1.0, 1.0), (-0.5, 3e9, 1.0),
            (1e-300, 1.0, 1.0), (-0.5, -1e-300, 1.0)]
    for q, (k, j, i) in enumerate(edge):
        p["k"][q], p["j"][q], p["i"][q] = k, j, i
    return p


def check(ctx, p, s_w_values, lat, lon, h):
    got = ctx.particle_records()
    src, dep, la, lo = po.particle_records(p, s_w_values, lat, lon, h)
    print(f"{len(p)} particles -> {len(src)} records", flush=True)
    assert len(got["id"]) == len(src)
    assert np.array_equal(got["id"], p["id"][src])
    for f in ("k", "j", "i", "health", "age", "time"):
        assert np.array_equal(got[f].view(np.uint64), p[f][src].view(np.uint64)), f
    assert np.array_equal(got["object_type"], p["driftObjectType"][src])     # Particles.cpp:144-145
    assert np.array_equal(got["drift_side"], p["driftSide"][src])
    for f, ref in (("depth", dep), ("lat", la), ("lon", lo)):
        assert np.array_equal(got[f].view(np.uint64), ref.view(np.uint64)), f
    return got


def test_roms_grid_records_match_oracle(ctx):
    c = random_roms(seed=21, s_rho=6, eta=40, xi=56)
    # a curvilinear, rotated grid (the ROMS path stores lat/lon/s_w for the records)
    jj, ii = np.meshgrid(np.arange(40.0), np.arange(56.0), indexing="ij")
    c["lon"] = 14 + 0.01 * (ii * np.cos(0.3) - jj * np.sin(0.3)) + 1e-4 * ii * jj
    c["lat"] = 40.5 + 0.01 * (ii * np.sin(0.3) + jj * np.cos(0.3)) - 2e-5 * ii * ii
    ctx.set_roms_grid(c["s_w"], c["lon"], c["lat"], c["h"], c["mask_rho"], c["mask_u"], c["mask_v"],
                      angle=c["angle"])
    p = positions(20_000, 7, 40, 56, seed=1)
    ctx.set_particles(p)
    h_land0 = np.where(c["mask_rho"] == 0, 0.0, c["h"])            # the adapter's h (ROMSAdapter.cpp:244-252)
    got = check(ctx, p, c["s_w"], c["lat"], c["lon"], h_land0)
    assert 0 < len(got["id"]) < len(p)


def test_full_size_records_match_oracle(ctx):
    W = full_window(T=2)
    ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask)
    s_w_values = np.linspace(-1.0, 0.0, W.s_w)
    lat, lon = np.degrees(W.latRad), np.degrees(W.lonRad)
    ctx.set_geo(lat, lon, s_w_values)
    p = positions(4_000_000, W.s_w, W.eta, W.xi, seed=2)
    ctx.set_particles(p)
    check(ctx, p, s_w_values, lat, lon, W.h)


def test_records_need_geo(ctx):
    W = full_window(T=2, eta=32, xi=24)
    ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask)
    ctx.set_particles(wc.make_particles(4, k=-0.5, j=1.0, i=1.0))
    with pytest.raises(wc.WcError):
        ctx.particle_records()


def test_restart_round_trip_through_device(ctx, tmp_path):
    W = full_window(T=2, eta=96, xi=80)
    ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask)
    s_w_values = np.linspace(-1.0, 0.0, W.s_w)
    ctx.set_geo(np.degrees(W.latRad), np.degrees(W.lonRad), s_w_values)
    p = positions(50_000, W.s_w, W.eta, W.xi, seed=3)
    ctx.set_particles(p)
    rec = ctx.particle_records()
    path = tmp_path / "restart.nc"
    wc.write_restart_nc(path, rec, 7200.0, 5489, ocean_model="roms")
    back, cp = wc.read_restart_nc(path, wc.default_config(), "roms")
    assert cp == 7200.0
    ctx.set_particles(back)
    again = ctx.get_particles()
    src = po.particle_records(p, s_w_values, np.degrees(W.latRad), np.degrees(W.lonRad), W.h)[0]
    assert np.array_equal(again["id"], p["id"][src])
    for f in ("k", "j", "i", "health", "age", "time"):
        assert np.array_equal(again[f].view(np.uint64), p[f][src].view(np.uint64)), f
    assert np.array_equal(again["driftObjectType"], p["driftObjectType"][src])
    assert np.array_equal(again["driftSide"], p["driftSide"][src])

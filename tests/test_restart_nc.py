"""Particle history/restart records and the restart file (host code, no device).

1. The oracle's kji2deplatlon / netcdfRecords (OceanModelAdapter.cpp:271-321,
   Particles.cpp:115-147) pinned by the fixtures of tests/VerticalCoordinateTest.cpp:
   on the rectilinear grid lat/lon are linear in j/i; h of the grids is bilinear,
   so depth is -h|sW(L)| at W level L and linear in k between levels; the
   curvilinear grid's lon is bilinear in (i, j) and reproduced exactly, its lat
   (with an i^2 term) interpolates i^2 linearly across each cell; the reference's
   range checks give the 1e37 sentinel; records keep input order.
2. wc_write_restart_nc: the file read by an independent classic-format parser
   (tests/ncparse.py) has the reference's attributes, dimensions and variables
   (Particles.cpp:150-264) and every value bit for bit.
3. wc_read_restart_nc: files written by scipy's NetCDF writer (CDF-1 and CDF-2,
   with and without a record dimension) and by wc_write_restart_nc, and the
   reference's rejections with its messages (Particles.cpp:378-432).
"""
import math

import numpy as np
import pytest
from scipy.io import netcdf_file

import ncparse
import pyoracle as po
import wacommplusplus_amd as wc

SW5 = np.array([-1.0, -0.6, -0.3, -0.1, 0.0])      # VerticalCoordinateTest.cpp:75-76


def rect_grid():
    """VerticalCoordinateTest.cpp:82-89: 3x3, lon 10+0.1i, lat 40+0.1j, h 80+20i+10j."""
    j, i = np.meshgrid(np.arange(3.0), np.arange(3.0), indexing="ij")
    return 40 + 0.1 * j, 10 + 0.1 * i, 80 + 20 * i + 10 * j


def curv_grid():
    """VerticalCoordinateTest.cpp:33-47: a rotated, sheared 5x6 grid."""
    th = 0.5
    j, i = np.meshgrid(np.arange(5.0), np.arange(6.0), indexing="ij")
    lon = 14 + 0.05 * (i * math.cos(th) - j * math.sin(th)) + 0.004 * i * j
    lat = 40.5 + 0.05 * (i * math.sin(th) + j * math.cos(th)) - 0.003 * i * i
    return lat, lon, 50 + 7 * i + 3 * j


def test_rectilinear_lat_lon_linear_and_depth_at_levels():
    lat, lon, h = rect_grid()
    rng = np.random.default_rng(3)
    for _ in range(200):
        j, i = rng.uniform(0, 2, 2) * (1 - 1e-9)
        hh = 80 + 20 * i + 10 * j
        for L in range(-3, 1):                              # valid levels -s_w+2 .. 0
            d, la, lo = po.kji2deplatlon(float(L), j, i, SW5, lat, lon, h)
            assert abs(la - (40 + 0.1 * j)) < 1e-12 and abs(lo - (10 + 0.1 * i)) < 1e-12
            assert abs(d + hh * abs(SW5[L + 4])) < 1e-9
        # linear in k between W levels
        for L in range(-3, 1):
            kF = rng.uniform(0, 1)
            d_up = po.kji2deplatlon(float(L), j, i, SW5, lat, lon, h)[0]
            d_lo = -hh * abs(SW5[L - 1 + 4])
            d = po.kji2deplatlon(L - kF, j, i, SW5, lat, lon, h)[0]
            assert abs(d - ((1 - kF) * d_up + kF * d_lo)) < 1e-9
    # the surface (VerticalCoordinateTest.cpp:100-102)
    assert po.kji2deplatlon(0.0, 1.3, 0.7, SW5, lat, lon, h)[0] == 0.0


def test_curvilinear_bilinear_fields():
    lat, lon, h = curv_grid()
    th = 0.5
    rng = np.random.default_rng(4)
    for _ in range(300):
        j, i = rng.uniform(0, 4, 1)[0] * (1 - 1e-9), rng.uniform(0, 5, 1)[0] * (1 - 1e-9)
        iI = math.floor(i)
        i2 = iI * iI + (2 * iI + 1) * (i - iI)                # i^2 interpolated linearly over the cell
        d, la, lo = po.kji2deplatlon(-0.37, j, i, SW5, lat, lon, h)
        assert abs(lo - (14 + 0.05 * (i * math.cos(th) - j * math.sin(th)) + 0.004 * i * j)) < 1e-12
        assert abs(la - (40.5 + 0.05 * (i * math.sin(th) + j * math.cos(th)) - 0.003 * i2)) < 1e-12
        hh = 50 + 7 * i + 3 * j
        assert abs(d - (-(hh * 0.0 + 0.37 * hh * 0.1))) < 1e-9      # k=-0.37 between sW 0 and -0.1


def test_sentinels_follow_reference_range_checks():
    lat, lon, h = curv_grid()
    bad = 1e37
    cases = [
        (-0.5, 4.0, 2.0),          # north node line: cell j+1 off the grid
        (-0.5, 2.0, 5.0),          # east node line
        (-0.5, 4.0, 5.0),
        (-4.0, 1.0, 1.0),          # ceil(k) = -s_w+1
        (-7.5, 1.0, 1.0),
        (0.5, 1.0, 1.0),           # ceil(k) = 1
        (-0.5, float("nan"), 1.0),
        (-0.5, 1.0, float("inf")),
        (float("nan"), 1.0, 1.0),
        (-0.5, 3e9, 1.0),          # beyond int range
    ]
    for k, j, i in cases:
        assert po.kji2deplatlon(k, j, i, SW5, lat, lon, h) == (bad, bad, bad), (k, j, i)
    # the south and west node lines and the lowest valid level are inside
    for k, j, i in [(-0.5, 0.0, 2.0), (-0.5, 2.0, 0.0), (-3.0, 1.0, 1.0), (-3.999, 1.0, 1.0), (0.0, 3.999, 4.999)]:
        assert po.kji2deplatlon(k, j, i, SW5, lat, lon, h)[0] != bad, (k, j, i)


def test_records_filter_and_order():
    lat, lon, h = curv_grid()
    keep = [(-0.2, 1.5, 2.5), (0.0, 0.0, 0.0), (-3.5, 3.9, 4.9), (-1.0, 2.0, 3.0)]
    drop = [(-0.2, -0.1, 2.0), (-0.2, 1.0, -1e-9), (0.2, 1.0, 1.0), (-0.2, 4.0, 1.0), (-4.0, 1.0, 1.0)]
    order = [("k", 0), ("d", 0), ("k", 1), ("d", 1), ("d", 2), ("k", 2), ("d", 3), ("d", 4), ("k", 3)]
    pos = [keep[q] if t == "k" else drop[q] for t, q in order]
    p = wc.make_particles(len(pos))
    p["k"], p["j"], p["i"] = np.array(pos).T
    src, dep, la, lo = po.particle_records(p, SW5, lat, lon, h)
    assert list(src) == [q for q, (t, _) in enumerate(order) if t == "k"]
    for r, q in enumerate(src):
        assert (dep[r], la[r], lo[r]) == po.kji2deplatlon(*pos[q], SW5, lat, lon, h)


# ---- the restart file --------------------------------------------------------

def records(n, seed=0):
    rng = np.random.default_rng(seed)
    r = {f: rng.normal(size=n) for f in wc.RECORD_FIELDS if f != "id"}
    r["id"] = rng.integers(0, 2 ** 63, n, dtype=np.uint64) * 2 + 1            # needs all 64 bits
    r["k"][:3] = [-0.0, 5e-324, np.finfo(float).max][:min(3, n)] if n >= 3 else r["k"][:3]
    return r


PROV = {"wacomm_git_revision": "abc123", "wacomm_compiler": "hipcc", "wacomm_cmake_options": "",
        "wacomm_configuration_file": "config.json", "wacomm_configuration": "{}"}


def test_restart_file_structure_and_values(tmp_path):
    n = 37
    r = records(n)
    ot = np.arange(n, dtype=np.uint32) % 5
    ds = (np.arange(n, dtype=np.int32) % 3) - 1
    path = tmp_path / "restart.nc"
    wc.write_restart_nc(path, r, 7200.5, 2 ** 64 - 3, backward=True, ocean_model="roms", attributes=PROV,
                        object_type=ot, drift_side=ds)
    assert not (tmp_path / "restart.nc.partial").exists()
    f = ncparse.read(path)
    assert f["version"] == 5 and f["numrecs"] == 0
    assert f["dims"] == [("particles", n), ("particle_time", 1)]
    g = f["gatts"]
    assert list(g) == list(PROV) + ["wacomm_random_key_scheme", "wacomm_restart_version", "tracking_direction",
                                    "ocean_model", "checkpoint_time", "random_seed"]
    assert g["wacomm_random_key_scheme"] == "ordered-splitmix64-v2" and g["wacomm_restart_version"] == "3"
    assert g["tracking_direction"] == "backward" and g["ocean_model"] == "roms"
    assert g["checkpoint_time"].dtype == ">f8" and g["checkpoint_time"][0] == 7200.5
    assert g["random_seed"].dtype == ">u8" and int(g["random_seed"][0]) == 2 ** 64 - 3
    v = f["variables"]
    assert list(v) == ["particle_time", "id", "object_type", "drift_side", "lat", "lon", "depth", "k", "j", "i",
                       "health", "age", "time"]
    assert v["particle_time"]["data"][0] == 7200.5 and v["particle_time"]["dims"] == ["particle_time"]
    assert v["particle_time"]["atts"]["units"] == "seconds since 1968-05-23 00:00:00 GMT"
    assert v["id"]["type"] == 11 and v["id"]["atts"] == {"long_name": "id", "cf_role": "trajectory_id"}
    assert np.array_equal(v["id"]["data"].astype(np.uint64), r["id"])
    assert v["object_type"]["type"] == 9 and np.array_equal(v["object_type"]["data"], ot)
    assert v["drift_side"]["type"] == 4 and np.array_equal(v["drift_side"]["data"], ds)
    for name in ("lat", "lon", "depth", "k", "j", "i", "health", "age", "time"):
        assert v[name]["dims"] == ["particles", "particle_time"] and v[name]["type"] == 6
        assert np.array_equal(v[name]["data"][:, 0].astype(np.float64).view(np.uint64),
                              r[name].view(np.uint64)), name
    assert v["depth"]["atts"] == {"long_name": "particle depth below sea surface", "unit": "meters",
                                  "standard_name": "depth", "positive": "up", "axis": "Z"}
    assert v["health"]["atts"]["units"] == "1e-9"
    assert v["time"]["atts"]["coordinates"] == "ocean_time lon lat depth"


def test_restart_round_trip_through_reader(tmp_path):
    n = 1000
    r = records(n, seed=2)
    path = tmp_path / "r.nc"
    wc.write_restart_nc(path, r, 3600.0, 5489, attributes=PROV)
    cfg = wc.default_config()
    p, cp = wc.read_restart_nc(path, cfg, "roms")
    assert cp == 3600.0 and len(p) == n
    assert np.array_equal(p["id"], r["id"])
    for f in ("k", "j", "i", "health", "age", "time"):
        assert np.array_equal(p[f].view(np.uint64), r[f].view(np.uint64)), f
    assert np.all(p["driftObjectType"] == 0) and np.all(p["driftSide"] == 0)


def test_empty_history_keeps_zero_length_variables(tmp_path):
    path = tmp_path / "empty.nc"
    r = {f: np.zeros(0, np.uint64 if f == "id" else float) for f in wc.RECORD_FIELDS}
    wc.write_restart_nc(path, r, 0.0, 5489)
    f = ncparse.read(path)
    assert f["dims"][0] == ("particles", 0) and f["variables"]["id"]["record"]
    assert f["variables"]["lat"]["data"].shape == (0, 1)
    p, cp = wc.read_restart_nc(path, wc.default_config(), "roms")
    assert len(p) == 0 and cp == 0.0


def scipy_restart(path, n, version, record, atts):
    rng = np.random.default_rng(n)
    with netcdf_file(path, "w", version=version) as f:
        for k, v in atts.items():
            setattr(f, k, v)
        f.createDimension("particles", None if record else n)
        f.createDimension("particle_time", 1)
        v = f.createVariable("id", "i4", ("particles",))
        v[:] = np.arange(n) * 3 + 1
        vals = {}
        for name in ("k", "j", "i", "health", "age", "time"):
            var = f.createVariable(name, "f8", ("particles", "particle_time"))
            vals[name] = rng.normal(size=(n, 1))
            var[:] = vals[name]
        ds = f.createVariable("drift_side", "i4", ("particles",))
        ds[:] = np.full(n, -1, np.int32)
    return vals


@pytest.mark.parametrize("version,record", [(1, False), (2, False), (1, True), (2, True)])
def test_reads_scipy_written_classic_files(tmp_path, version, record):
    path = tmp_path / "s.nc"
    vals = scipy_restart(path, 23, version, record,
                         {"wacomm_restart_version": "3", "tracking_direction": "forward",
                          "checkpoint_time": np.float64(1800.0), "ocean_model": "roms"})
    p, cp = wc.read_restart_nc(path, wc.default_config(), "roms")
    assert cp == 1800.0
    assert np.array_equal(p["id"], np.arange(23) * 3 + 1)
    for name, a in vals.items():
        assert np.array_equal(p[name], a[:, 0]), name
    assert np.all(p["driftSide"] == -1) and np.all(p["driftObjectType"] == 0)


def test_file_without_version_is_not_checked(tmp_path):
    path = tmp_path / "old.nc"
    scipy_restart(path, 5, 1, False, {"tracking_direction": "backward"})
    p, cp = wc.read_restart_nc(path, wc.default_config(), "wrf")
    assert len(p) == 5 and cp is None


@pytest.mark.parametrize("atts,cfg_kw,model,msg", [
    ({"wacomm_restart_version": "4"}, {}, "roms", "Unsupported restart version: 4"),
    ({"wacomm_restart_version": "2", "tracking_direction": "forward"}, {"driftModel": 1}, "roms",
     "Restart version 2 does not contain drift object state required by drift.model=leeway"),
    ({"wacomm_restart_version": "3", "tracking_direction": "backward"}, {}, "roms",
     "Restart tracking direction is backward but configuration direction is forward"),
    ({"wacomm_restart_version": "3", "tracking_direction": "forward"}, {"trackingDirection": -1}, "roms",
     "Restart tracking direction is forward but configuration direction is backward"),
    ({"wacomm_restart_version": "3", "tracking_direction": "forward", "random_seed": np.int32(7)}, {}, "roms",
     "Restart random seed does not match the configured random seed"),
    ({"wacomm_restart_version": "3", "tracking_direction": "forward", "wacomm_random_key_scheme": "legacy"}, {},
     "roms", "Restart random-key scheme is legacy but this build uses ordered-splitmix64-v2"),
    ({"wacomm_restart_version": "3", "tracking_direction": "forward", "ocean_model": "nemo"}, {}, "roms",
     "Restart ocean model is nemo but configuration ocean model is roms"),
])
def test_reference_rejections(tmp_path, atts, cfg_kw, model, msg):
    path = tmp_path / "bad.nc"
    scipy_restart(path, 4, 2, False, atts)
    with pytest.raises(wc.WcError) as e:
        wc.read_restart_nc(path, wc.default_config(**cfg_kw), model)
    assert e.value.status == wc.WC_ERR_INVALID and msg in str(e.value)


def test_seed_and_version_two_accepted(tmp_path):
    path = tmp_path / "v2.nc"
    scipy_restart(path, 4, 2, False, {"wacomm_restart_version": "2", "tracking_direction": "forward",
                                      "random_seed": np.int32(5489)})
    p, _ = wc.read_restart_nc(path, wc.default_config(), "roms")
    assert len(p) == 4


def test_unreadable_files(tmp_path):
    h5 = tmp_path / "nc4.nc"
    h5.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))
    with pytest.raises(wc.WcError) as e:
        wc.read_restart_nc(h5, wc.default_config(), "roms")
    assert e.value.status == wc.WC_ERR_UNSUPPORTED
    junk = tmp_path / "junk.nc"
    junk.write_bytes(b"CDF\x05" + bytes(6))
    with pytest.raises(wc.WcError) as e:
        wc.read_restart_nc(junk, wc.default_config(), "roms")
    assert e.value.status == wc.WC_ERR_IO
    with pytest.raises(wc.WcError) as e:
        wc.read_restart_nc(tmp_path / "missing.nc", wc.default_config(), "roms")
    assert e.value.status == wc.WC_ERR_IO


def test_writer_rejects_repeated_attribute(tmp_path):
    r = records(3)
    with pytest.raises(wc.WcError):
        wc.write_restart_nc(tmp_path / "x.nc", r, 0.0, 1, attributes={"random_seed": "1"})


FUZZ = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import wacommplusplus_amd as wc
src, n_cases, seed = sys.argv[2], int(sys.argv[3]), int(sys.argv[4])
good = open(src, "rb").read()
rng = np.random.default_rng(seed)
cfg = wc.default_config(randomSeed=5489)
path = src + ".fuzz"
accepted = rejected = 0
for case in range(n_cases):
    b = bytearray(good)
    kind = case % 4
    if kind == 0:                                   # flip bytes in and just after the header
        for _ in range(int(rng.integers(1, 5))):
            b[int(rng.integers(4, min(len(b), 1400)))] ^= 1 << int(rng.integers(8))
    elif kind == 1:                                 # truncate
        del b[int(rng.integers(0, len(b))):]
    elif kind == 2:                                 # a header word set to a huge or random value
        at = int(rng.integers(1, min(len(b), 1400) // 4)) * 4
        b[at:at + 4] = bytes([0xFF] * 4) if rng.integers(2) else rng.integers(0, 256, 4, dtype=np.uint8).tobytes()
    else:                                           # an 8-byte field set to a huge or random value
        at = int(rng.integers(1, min(len(b), 1400) // 8)) * 8
        pick = int(rng.integers(3))
        b[at:at + 8] = (bytes([0x7F] + [0xFF] * 7) if pick == 0 else bytes([0xFF] * 7 + [0xF8]) if pick == 1
                        else rng.integers(0, 256, 8, dtype=np.uint8).tobytes())
    open(path, "wb").write(bytes(b))
    try:
        p, _ = wc.read_restart_nc(path, cfg, "roms")
        assert len(p) * 8 <= len(b), (case, len(p), len(b))
        accepted += 1
    except wc.WcError:
        rejected += 1
print("ok", accepted, rejected, flush=True)
"""


def test_reader_survives_corrupt_files(tmp_path):
    """A restart file from disk is untrusted input: byte flips, truncations
    and huge header values give a clean error or a bounded population, never
    an out-of-bounds read.  Run in a child process, so that a crash fails the
    test instead of the test session."""
    import os
    import subprocess
    import sys
    src = tmp_path / "good.nc"
    wc.write_restart_nc(src, records(40, seed=9), 3600.0, 5489, ocean_model="roms", attributes=PROV)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", FUZZ, root, str(src), "2000", "7"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    ok, accepted, rejected = r.stdout.split()
    assert ok == "ok" and int(accepted) + int(rejected) == 2000
    assert int(rejected) > 0


def test_reader_under_address_sanitizer(tmp_path):
    """The reader's source built with g++ under AddressSanitizer and
    UndefinedBehaviorSanitizer (host code only), fed 4,000 corrupted restart
    files (tests/cpp/restart_fuzz.cpp): no invalid access, and a population
    reported only when it fits the file.  The reader before the overflow-safe
    offsets reported 4,099,880,709 particles from a 6,708-byte file here."""
    import os
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "restart_fuzz"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-I", os.path.join(root, "include"), "-x", "c++",
                    os.path.join(root, "wacommplusplus_amd", "csrc", "wc_netcdf.hip"), "-x", "none",
                    os.path.join(root, "tests", "cpp", "restart_fuzz.cpp"), "-o", str(exe)], check=True)
    src = tmp_path / "good.nc"
    wc.write_restart_nc(src, records(40, seed=9), 3600.0, 5489, ocean_model="roms", attributes=PROV)
    env = dict(os.environ, ASAN_OPTIONS="verify_asan_link_order=0:detect_leaks=1")
    r = subprocess.run([str(exe), str(src), "4000", "7", str(tmp_path / "case.nc")], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-1000:] + r.stderr[-3000:]
    ok, accepted, rejected = r.stdout.split()
    assert ok == "ok" and int(accepted) + int(rejected) == 4000 and int(rejected) > 0

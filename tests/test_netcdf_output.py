"""The concentration file of Wacomm::save (Wacomm.cpp:913-1020) written by
wc_write_concentration_nc, read back with scipy's independent NetCDF classic
reader: dimensions (including the reference's "eta_xi" name), variables,
attributes, _FillValue and every value, bit for bit.  Host code: no device."""
import numpy as np
import pytest
from scipy.io import netcdf_file

import wacommplusplus_amd as wc


def sample(T=3, s_rho=4, eta=5, xi=7, seed=1):
    rng = np.random.default_rng(seed)
    lat = 40 + np.arange(eta)[:, None] * 0.01 + 0 * np.arange(xi)[None]
    lon = 14 + 0 * np.arange(eta)[:, None] + np.arange(xi)[None] * 0.01
    mask = (rng.random((eta, xi)) > 0.3).astype(float)
    h = np.where(mask > 0, rng.uniform(10, 50, (eta, xi)), 0.0)
    zeta = rng.normal(0, 0.1, (T, eta, xi)).astype(np.float32)
    conc = rng.integers(0, 50, (T, s_rho, eta, xi)).astype(np.float32)
    conc[:, :, mask == 0] = np.float32(1e37)                       # land (Wacomm.cpp:724-739)
    return dict(ocean_time=np.arange(T) * 3600.0 + 1.6e9, s_rho_values=(np.arange(s_rho) - s_rho + 0.5) / s_rho,
                lat_rho=lat, lon_rho=lon, mask_rho=mask, h=h, zeta=zeta, conc=conc)


def test_round_trip(tmp_path):
    d = sample()
    path = tmp_path / "conc.nc"
    wc.write_concentration_nc(path, **d, attributes={"wacomm_random_key_scheme": "ordered-splitmix64-v2"})
    assert not (tmp_path / "conc.nc.partial").exists()
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.version_byte == 2
        assert f.dimensions == {"ocean_time": 3, "s_rho": 4, "eta_rho": 5, "eta_xi": 7}
        assert f.wacomm_random_key_scheme == b"ordered-splitmix64-v2"
        v = f.variables
        assert v["conc"].dimensions == ("ocean_time", "s_rho", "eta_rho", "eta_xi")
        assert v["zeta"].dimensions == ("ocean_time", "eta_rho", "eta_xi")
        assert v["lat_rho"].dimensions == ("eta_rho", "eta_xi")
        for name in ("ocean_time", "s_rho_values", "lat_rho", "lon_rho", "mask_rho", "h"):
            key = "s_rho" if name == "s_rho_values" else name
            assert v[key].data.dtype.kind == "f" and v[key].data.itemsize == 8
            assert np.array_equal(v[key].data, d[name]), name
        # stored big-endian: compare the bits of the native values
        assert np.array_equal(v["zeta"].data.astype(np.float32).view(np.uint32), d["zeta"].view(np.uint32))
        assert np.array_equal(v["conc"].data.astype(np.float32).view(np.uint32), d["conc"].view(np.uint32))
        assert v["conc"]._FillValue == np.float32(9.99999993e36)
        assert v["conc"].long_name == b"concentration_of_suspended_matter_in_sea_water"
        assert v["ocean_time"].units == b"seconds since 1968-05-23 00:00:00 GMT"
        assert v["s_rho"].standard_name == b"ocean_sigma_coordinates"
        assert v["h"].units == b"meter"


def test_unwritten_fields_are_fill(tmp_path):
    d = sample()
    d["conc"] = None
    d["zeta"] = None
    path = tmp_path / "empty.nc"
    wc.write_concentration_nc(path, **d)
    with netcdf_file(path, "r", mmap=False) as f:
        assert np.all(f.variables["conc"].data == np.float32(9.99999993e36))
        assert np.all(f.variables["zeta"].data == np.float32(9.99999993e36))


def test_unwritable_path_is_an_error(tmp_path):
    d = sample()
    with pytest.raises(wc.WcError) as e:
        wc.write_concentration_nc(tmp_path / "missing-dir" / "x.nc", **d)
    assert e.value.status == wc.WC_ERR_IO

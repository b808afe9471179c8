"""ctypes binding of libwacomm_hip.so (include/wacomm_hip.h).

The library is built in-tree by wacommplusplus_amd.build; loading fails loudly
when it is missing -- there is no CPU fallback in this package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .types import PARTICLE_DTYPE, PATH, STORE, ConfigData

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libwacomm_hip.so")

WC_OK, WC_ERR_HIP, WC_ERR_INVALID, WC_ERR_UNSUPPORTED, WC_ERR_NOMEM, WC_ERR_NO_DEVICE, WC_ERR_IO = range(7)
_STATUS = {1: "WC_ERR_HIP", 2: "WC_ERR_INVALID", 3: "WC_ERR_UNSUPPORTED", 4: "WC_ERR_NOMEM",
           5: "WC_ERR_NO_DEVICE", 6: "WC_ERR_IO"}

# every symbol include/wacomm_hip.h declares
EXPORTS = (
    "wc_hip_move_particles", "wc_hip_concentration", "wc_ctx_create", "wc_ctx_destroy", "wc_last_error",
    "wc_sync", "wc_set_grid", "wc_upload_record", "wc_set_particles", "wc_append_particles",
    "wc_get_particles", "wc_particle_count", "wc_move_interval", "wc_remove_dead", "wc_grid",
    "wc_set_timing", "wc_kernel_time", "wc_reset_kernel_time", "wc_step_count", "wc_reset_step_count",
    "wc_version", "wc_cell_metric_table", "wc_philox4x32_10", "wc_emit", "wc_run_window",
    "wc_set_roms_grid", "wc_upload_roms_record", "wc_get_record", "wc_check_roms_grid", "wc_compare_records",
    "wc_ens_set", "wc_ens_set_launch", "wc_ens_get_launch", "wc_ens_autotune", "wc_ens_run", "wc_ens_surface",
    "wc_ens_get", "wc_ens_release", "wc_ens_set_sort", "wc_ens_get_sort", "wc_ens_set_lds", "wc_ens_get_lds", "wc_unet_create", "wc_unet_destroy", "wc_unet_last_error", "wc_unet_param_count",
    "wc_unet_get_params", "wc_unet_set_params", "wc_unet_forward", "wc_unet_gradients", "wc_unet_train_step",
    "wc_unet_create_ex", "wc_unet_gradients_weighted", "wc_unet_train_step_weighted", "wc_unet_set_graph",
    "wc_write_concentration_nc", "wc_write_restart_nc", "wc_read_restart_nc", "wc_set_geo",
    "wc_particle_records", "wc_set_event_tracking", "wc_clear_events", "wc_get_events", "wc_grid_noalloc",
    "wc_set_sort", "wc_get_sort", "wc_autotune_sort", "wc_set_lds", "wc_get_lds",
)


class Source(C.Structure):
    """wc_source: release point and schedule (Source.hpp, SourceEmission.hpp:12-16)."""
    _fields_ = [("k", C.c_double), ("j", C.c_double), ("i", C.c_double), ("mode", C.c_int),
                ("value", C.c_double), ("start", C.c_double), ("end", C.c_double)]


UNIFORM_RATE, FORCING_INTERVAL_BATCH, SINGLE_PULSE = 0, 1, 2


class WindowDesc(C.Structure):
    _fields_ = [("T", C.c_int), ("oceanTime", C.c_void_p), ("zeta", C.c_void_p), ("u", C.c_void_p),
                ("v", C.c_void_p), ("w", C.c_void_p), ("akt", C.c_void_p)]


class RunStats(C.Structure):
    _fields_ = [("solver_s", C.c_double), ("move_ms", C.c_double), ("particle_steps", C.c_uint64),
                ("particles", C.c_size_t), ("intervals", C.c_int)]


ROMS_MAX_MISSING = 14      # WC_ROMS_MAX_MISSING


class RomsGridDesc(C.Structure):
    """wc_roms_grid: the static grid of a ROMS file (ROMSAdapter.cpp:56-166)."""
    _fields_ = [(n, C.c_int) for n in ("s_rho", "eta_rho", "xi_rho", "eta_u", "xi_u", "eta_v", "xi_v")] + \
               [(n, C.c_void_p) for n in ("s_w", "lon_rho", "lat_rho", "h", "mask_rho", "mask_u", "mask_v",
                                          "angle")] + \
               [("earth_relative", C.c_int)] + \
               [(n, C.c_void_p) for n in ("missing_u", "missing_v", "missing_angle")] + \
               [(n, C.c_int) for n in ("n_missing_u", "n_missing_v", "n_missing_angle")] + \
               [("s_rho_values", C.c_void_p)]


class ConcFile(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("T", "s_rho", "eta", "xi")] + \
               [(n, C.c_void_p) for n in ("ocean_time", "s_rho_values", "lat_rho", "lon_rho", "mask_rho", "h",
                                          "zeta", "conc", "global_attributes")]


class RestartFile(C.Structure):
    """wc_restart_file: a restart-version-3 particle file (Particles.cpp:150-264)."""
    _fields_ = [("n", C.c_size_t), ("particle_time", C.c_double), ("random_seed", C.c_uint64),
                ("backward", C.c_int), ("ocean_model", C.c_char_p), ("global_attributes", C.c_void_p),
                ("id", C.c_void_p), ("object_type", C.c_void_p), ("drift_side", C.c_void_p)] + \
               [(n, C.c_void_p) for n in ("lat", "lon", "depth", "k", "j", "i", "health", "age", "time")]


RECORD_FIELDS = ("id", "lat", "lon", "depth", "k", "j", "i", "health", "age", "time")


class Records(C.Structure):
    """wc_records: host arrays receiving Particles::netcdfRecords."""
    _fields_ = [(n, C.c_void_p) for n in RECORD_FIELDS] + [("object_type", C.c_void_p), ("drift_side", C.c_void_p)]


class EnsSource(C.Structure):
    """wc_ens_source: a release point of a device-generated ensemble."""
    _fields_ = [("k", C.c_double), ("j", C.c_double), ("i", C.c_double), ("spread", C.c_double),
                ("time", C.c_double), ("count", C.c_uint64), ("scenario", C.c_int), ("sub_scenario", C.c_int),
                ("sub_every", C.c_uint32)]


class EnsStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("wall_s", C.c_double), ("particle_steps", C.c_uint64),
                ("intervals", C.c_int)]


class WcError(RuntimeError):
    def __init__(self, status, message=""):
        self.status = status
        super().__init__(f"{_STATUS.get(status, status)}: {message}")


_lib = None

vp = C.c_void_p


def load_library(path: str = LIB_PATH):
    """Loads libwacomm_hip.so (once) and declares the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m wacommplusplus_amd.build` "
                          "(the HIP backend has no CPU fallback)")
    L = C.CDLL(path)
    i, d, sz = C.c_int, C.c_double, C.c_size_t
    sig = {
        "wc_hip_move_particles": [vp, vp, i, i, i, i, i, i] + [vp] * 6 + [vp] * 9 + [i, i, i, vp],
        "wc_hip_concentration": [vp, i, i, i, i, vp, vp],
        "wc_grid_noalloc": [vp, i, i, i, i, vp, i, vp, vp],
        "wc_ctx_create": [i, C.POINTER(vp)],
        "wc_ctx_destroy": [vp],
        "wc_sync": [vp],
        "wc_set_grid": [vp, i, i, i, i, vp, vp, vp, vp, vp, i, i],
        "wc_upload_record": [vp, i, vp, vp, vp, vp, vp],
        "wc_set_particles": [vp, vp, sz],
        "wc_append_particles": [vp, vp, sz],
        "wc_get_particles": [vp, vp, sz],
        "wc_particle_count": [vp, C.POINTER(sz)],
        "wc_set_event_tracking": [vp, i],
        "wc_set_sort": [vp, i, i],
        "wc_get_sort": [vp, C.POINTER(i), C.POINTER(i)],
        "wc_set_lds": [vp, i, i],
        "wc_get_lds": [vp, C.POINTER(i), C.POINTER(i)],
        "wc_autotune_sort": [vp, C.POINTER(ConfigData), d, d, i, i, i, C.POINTER(i), C.POINTER(d)],
        "wc_clear_events": [vp],
        "wc_get_events": [vp, vp, sz],
        "wc_move_interval": [vp, C.POINTER(ConfigData), d, d, i, i, i],
        "wc_remove_dead": [vp, C.POINTER(sz)],
        "wc_grid": [vp, vp, vp, i],
        "wc_set_timing": [vp, i],
        "wc_kernel_time": [vp, i, C.POINTER(d), C.POINTER(C.c_long)],
        "wc_reset_kernel_time": [vp],
        "wc_step_count": [vp, C.POINTER(C.c_uint64)],
        "wc_reset_step_count": [vp],
        "wc_emit": [vp, C.POINTER(ConfigData), vp, i, d, d, C.POINTER(sz)],
        "wc_run_window": [vp, C.POINTER(ConfigData), C.POINTER(WindowDesc), vp, i, i, vp, i, C.POINTER(RunStats)],
        "wc_set_roms_grid": [vp, C.POINTER(RomsGridDesc), i, i],
        "wc_upload_roms_record": [vp, i, vp, vp, vp, vp, vp],
        "wc_get_record": [vp, i, vp, vp, vp, vp, vp],
        "wc_check_roms_grid": [vp, C.POINTER(RomsGridDesc)],
        "wc_compare_records": [vp, i, i, C.POINTER(i), C.POINTER(C.c_longlong)],
        "wc_ens_set": [vp, vp, vp, sz, vp, i],
        "wc_ens_set_launch": [vp, i, i, i],
        "wc_ens_get_launch": [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i)],
        "wc_ens_autotune": [vp, C.POINTER(ConfigData), d, d, i, i, C.POINTER(d)],
        "wc_ens_run": [vp, C.POINTER(ConfigData), vp, vp, i, i, i, C.POINTER(EnsStats)],
        "wc_ens_surface": [vp, vp],
        "wc_ens_get": [vp, vp, vp, sz],
        "wc_ens_release": [vp, vp, i, vp, i, C.c_uint64],
        "wc_ens_set_sort": [vp, i, i],
        "wc_ens_get_sort": [vp, C.POINTER(i), C.POINTER(i)],
        "wc_ens_set_lds": [vp, i, i],
        "wc_ens_get_lds": [vp, C.POINTER(i), C.POINTER(i)],
        "wc_set_geo": [vp, vp, vp, vp],
        "wc_particle_records": [vp, C.POINTER(Records), sz, C.POINTER(sz)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.wc_last_error.argtypes = [vp]
    L.wc_last_error.restype = C.c_char_p
    L.wc_version.argtypes = []
    L.wc_version.restype = C.c_char_p
    L.wc_cell_metric_table.argtypes = [i, i, vp, vp, vp, vp]
    L.wc_cell_metric_table.restype = None
    L.wc_philox4x32_10.argtypes = [vp, vp, vp]
    L.wc_philox4x32_10.restype = None
    unet = {
        "wc_unet_create": [i, C.c_uint64, C.POINTER(vp)],
        "wc_unet_create_ex": [i, C.c_uint64, i, C.POINTER(vp)],
        "wc_unet_destroy": [vp],
        "wc_unet_get_params": [vp, vp],
        "wc_unet_set_params": [vp, vp],
        "wc_unet_forward": [vp, vp, i, i, i, vp, C.POINTER(d)],
        "wc_unet_gradients": [vp, vp, vp, i, i, i, vp, C.POINTER(C.c_float)],
        "wc_unet_train_step": [vp, vp, vp, i, i, i, C.c_float, C.POINTER(C.c_float)],
        "wc_unet_gradients_weighted": [vp, vp, vp, vp, i, i, i, vp, C.POINTER(C.c_float)],
        "wc_unet_train_step_weighted": [vp, vp, vp, vp, i, i, i, C.c_float, C.POINTER(C.c_float)],
        "wc_unet_set_graph": [vp, i],
    }
    for name, args in unet.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.wc_unet_last_error.argtypes = [vp]
    L.wc_unet_last_error.restype = C.c_char_p
    L.wc_unet_param_count.argtypes = [vp]
    L.wc_unet_param_count.restype = C.c_size_t
    L.wc_write_concentration_nc.argtypes = [C.c_char_p, C.POINTER(ConcFile)]
    L.wc_write_concentration_nc.restype = C.c_int
    L.wc_write_restart_nc.argtypes = [C.c_char_p, C.POINTER(RestartFile)]
    L.wc_write_restart_nc.restype = C.c_int
    L.wc_read_restart_nc.argtypes = [C.c_char_p, C.POINTER(ConfigData), C.c_char_p, vp, sz, C.POINTER(sz),
                                     C.POINTER(d), C.c_char_p, sz]
    L.wc_read_restart_nc.restype = C.c_int
    _lib = L
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(vp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Context:
    """One device context: the particle population (SoA on the device), a
    ring of forcing records and the count frame.  Mirrors the roles of
    Wacomm::run's device state (Wacomm.cpp:386-610) without per-interval
    re-allocation."""

    def __init__(self, device: int = 0):
        self.L = load_library()
        h = vp()
        rc = self.L.wc_ctx_create(device, C.byref(h))
        if rc:
            raise WcError(rc, f"wc_ctx_create(device={device})")
        self.h = h
        self.device = device
        self.shape = None

    def _chk(self, rc):
        if rc:
            raise WcError(rc, self.L.wc_last_error(self.h).decode(errors="replace"))

    def close(self):
        if self.h:
            self.L.wc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- grid and forcing ---------------------------------------------------
    def set_grid(self, lonRad, latRad, h, depthIntervals, s_rho, mask=None, n_slots=2, store="f32"):
        lon, lat, hh, dep = _f64(lonRad), _f64(latRad), _f64(h), _f64(depthIntervals)
        eta, xi = hh.shape
        m = None if mask is None else _f64(mask)
        self._keep = (lon, lat, hh, dep, m)
        self._chk(self.L.wc_set_grid(self.h, s_rho + 1, s_rho, eta, xi, _p(lon), _p(lat), _p(hh),
                                     None if m is None else _p(m), _p(dep), n_slots, STORE[store]))
        self.shape = (s_rho, eta, xi)

    def upload_record(self, slot, zeta, u, v, w, akt):
        arrs = [_f32(a) for a in (zeta, u, v, w, akt)]
        s_rho, eta, xi = self.shape
        assert arrs[0].shape == (eta, xi) and arrs[1].shape == (s_rho, eta, xi) == arrs[2].shape
        assert arrs[3].shape == (s_rho + 1, eta, xi) == arrs[4].shape
        self._chk(self.L.wc_upload_record(self.h, slot, *[_p(a) for a in arrs]))

    @staticmethod
    def _roms_desc(s_w, lon_rho, lat_rho, h, mask_rho, mask_u, mask_v, angle, earth_relative, missing_u,
                   missing_v, missing_angle, s_rho_values):
        sw, lon, lat, hh, mr, mu, mv = (_f64(a) for a in (s_w, lon_rho, lat_rho, h, mask_rho, mask_u, mask_v))
        ang = None if angle is None else _f64(angle)
        sr = None if s_rho_values is None else _f64(s_rho_values)
        miss = [_f64(np.r_[m, 0.0]) for m in (missing_u, missing_v, missing_angle)]
        eta, xi = mr.shape
        desc = RomsGridDesc(len(sw) - 1, eta, xi, *mu.shape, *mv.shape,
                            *[_p(a).value for a in (sw, lon, lat, hh, mr, mu, mv)],
                            None if ang is None else _p(ang).value, int(bool(earth_relative)),
                            *[_p(m).value for m in miss], len(missing_u), len(missing_v), len(missing_angle),
                            None if sr is None else _p(sr).value)
        return desc, (sw, lon, lat, hh, mr, mu, mv, ang, miss, sr)

    def set_roms_grid(self, s_w, lon_rho, lat_rho, h, mask_rho, mask_u, mask_v, angle=None,
                      earth_relative=False, missing_u=(), missing_v=(), missing_angle=(), s_rho_values=None,
                      n_slots=2, store="f32"):
        """The static part of ROMSAdapter::process (ROMSAdapter.cpp:146-151,
        228-275): lon/lat in degrees, s_w bottom first, masks 0/1, angle in
        radians (unused when earth_relative).  Leaves a grid for which
        upload_roms_record normalises raw ROMS records."""
        desc, keep = self._roms_desc(s_w, lon_rho, lat_rho, h, mask_rho, mask_u, mask_v, angle, earth_relative,
                                     missing_u, missing_v, missing_angle, s_rho_values)
        self._keep = keep
        self._chk(self.L.wc_set_roms_grid(self.h, C.byref(desc), n_slots, STORE[store]))
        mr, mu, mv = keep[4], keep[5], keep[6]
        self.shape = (len(keep[0]) - 1,) + mr.shape
        self.roms_shape = (mu.shape, mv.shape)

    def check_roms_grid(self, s_w, lon_rho, lat_rho, h, mask_rho, mask_u, mask_v, s_rho_values=None):
        """appendBoundaryRecord's grid checks (OceanModelAdapter.cpp:199-214)
        of the next file against the loaded ROMS grid; raises WcError."""
        desc, keep = self._roms_desc(s_w, lon_rho, lat_rho, h, mask_rho, mask_u, mask_v, None, True, (), (), (),
                                     s_rho_values)
        self._chk(self.L.wc_check_roms_grid(self.h, C.byref(desc)))

    def compare_records(self, slot_a, slot_b):
        """Duplicate-boundary check (OceanModelAdapter.cpp:216-233):
        (conflict, point) with conflict 0 none, 1 zeta, 2 u/v, 3 w/AKt."""
        conflict, point = C.c_int(), C.c_longlong()
        self._chk(self.L.wc_compare_records(self.h, slot_a, slot_b, C.byref(conflict), C.byref(point)))
        return conflict.value, point.value

    def upload_roms_record(self, slot, zeta, u, v, w, akt):
        """One raw ROMS record (u on the u grid, v on the v grid) normalised
        into ring slot `slot` on the device."""
        arrs = [_f32(a) for a in (zeta, u, v, w, akt)]
        s_rho, eta, xi = self.shape
        (ush, vsh) = self.roms_shape
        assert arrs[0].shape == (eta, xi) and arrs[1].shape == (s_rho,) + ush and arrs[2].shape == (s_rho,) + vsh
        assert arrs[3].shape == (s_rho + 1, eta, xi) == arrs[4].shape
        self._chk(self.L.wc_upload_roms_record(self.h, slot, *[_p(a) for a in arrs]))

    def get_record(self, slot):
        """Ring slot `slot` as (zeta, u, v, w, akt) FP32 host arrays."""
        s_rho, eta, xi = self.shape
        out = (np.zeros((eta, xi), np.float32), np.zeros((s_rho, eta, xi), np.float32),
               np.zeros((s_rho, eta, xi), np.float32), np.zeros((s_rho + 1, eta, xi), np.float32),
               np.zeros((s_rho + 1, eta, xi), np.float32))
        self._chk(self.L.wc_get_record(self.h, slot, *[_p(a) for a in out]))
        return out

    # -- particles ----------------------------------------------------------
    def set_particles(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=PARTICLE_DTYPE)
        self._chk(self.L.wc_set_particles(self.h, _p(p), len(p)))

    def append_particles(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=PARTICLE_DTYPE)
        self._chk(self.L.wc_append_particles(self.h, _p(p), len(p)))

    def count(self) -> int:
        n = C.c_size_t()
        self._chk(self.L.wc_particle_count(self.h, C.byref(n)))
        return n.value

    def get_particles(self) -> np.ndarray:
        n = self.count()
        out = np.zeros(n, dtype=PARTICLE_DTYPE)
        self._chk(self.L.wc_get_particles(self.h, _p(out), n))
        return out

    # -- cell-tile order of the moves ----------------------------------------
    def set_lds(self, chunk: int, halo: int = 2):
        """LDS-staged forcing for the mixed-precision move (wc_set_lds);
        chunk 0 turns it off.  Results are those of the plain move."""
        self._chk(self.L.wc_set_lds(self.h, int(chunk), int(halo)))

    def get_lds(self):
        c, h = C.c_int(), C.c_int()
        self._chk(self.L.wc_get_lds(self.h, C.byref(c), C.byref(h)))
        return c.value, h.value

    def set_sort(self, tile: int, period: int = 1):
        self._chk(self.L.wc_set_sort(self.h, int(tile), int(period)))

    def get_sort(self):
        t, p = C.c_int(), C.c_int()
        self._chk(self.L.wc_get_sort(self.h, C.byref(t), C.byref(p)))
        return t.value, p.value

    def autotune_sort(self, config: ConfigData, start: float, end: float, slot0: int, slot1: int, path="parity"):
        t, ms = C.c_int(), C.c_double()
        self._chk(self.L.wc_autotune_sort(self.h, C.byref(config), start, end, slot0, slot1, PATH[path],
                                          C.byref(t), C.byref(ms)))
        return t.value, ms.value

    # -- closure events (WC_EVENT_* bits per particle) ----------------------
    def set_event_tracking(self, on: bool = True):
        self._chk(self.L.wc_set_event_tracking(self.h, int(bool(on))))

    def clear_events(self):
        self._chk(self.L.wc_clear_events(self.h))

    def get_events(self) -> np.ndarray:
        n = self.count()
        out = np.zeros(n, dtype=np.uint8)
        self._chk(self.L.wc_get_events(self.h, _p(out), n))
        return out

    def set_geo(self, lat, lon, s_w_values):
        """The grid's lat, lon (degrees, [eta][xi]) and s_w values (bottom first)
        used by particle_records."""
        la, lo, sw = _f64(lat), _f64(lon), _f64(s_w_values)
        self._chk(self.L.wc_set_geo(self.h, _p(la), _p(lo), _p(sw)))

    def particle_records(self) -> dict:
        """Particles::netcdfRecords (Particles.cpp:115-147) computed on the
        device: a dict of the RECORD_FIELDS arrays plus object_type (uint32)
        and drift_side (int32)."""
        n = C.c_size_t()
        self._chk(self.L.wc_particle_records(self.h, None, 0, C.byref(n)))
        out = {f: np.zeros(n.value, np.uint64 if f == "id" else np.float64) for f in RECORD_FIELDS}
        out["object_type"] = np.zeros(n.value, np.uint32)
        out["drift_side"] = np.zeros(n.value, np.int32)
        r = Records(*[_p(out[f]).value for f in RECORD_FIELDS + ("object_type", "drift_side")])
        m = C.c_size_t()
        self._chk(self.L.wc_particle_records(self.h, C.byref(r), n.value, C.byref(m)))
        assert m.value == n.value
        return out

    # -- the hot path -------------------------------------------------------
    def move_interval(self, config: ConfigData, start: float, end: float, slot0: int, slot1: int,
                      path: str = "parity"):
        self._chk(self.L.wc_move_interval(self.h, C.byref(config), float(start), float(end), slot0, slot1,
                                          PATH[path]))

    def remove_dead(self) -> int:
        n = C.c_size_t()
        self._chk(self.L.wc_remove_dead(self.h, C.byref(n)))
        return n.value

    def grid(self, counts: bool = True, conc: bool = False, mask_output: bool = False):
        s_rho, eta, xi = self.shape
        c = np.zeros((s_rho, eta, xi), np.uint32) if counts else None
        f = np.zeros((s_rho, eta, xi), np.float32) if conc else None
        self._chk(self.L.wc_grid(self.h, None if c is None else _p(c), None if f is None else _p(f),
                                 int(mask_output)))
        if counts and conc:
            return c, f
        return c if counts else f

    def sync(self):
        self._chk(self.L.wc_sync(self.h))

    def emit(self, config: ConfigData, sources, start: float, end: float) -> int:
        """Sources::emit for one interval (Sources.cpp:233-248)."""
        arr = (Source * max(1, len(sources)))(*sources)
        n = C.c_size_t()
        self._chk(self.L.wc_emit(self.h, C.byref(config), arr, len(sources), float(start), float(end), C.byref(n)))
        return n.value

    def run_window(self, config: ConfigData, times, zeta, u, v, w, akt, sources=(), path="parity",
                   conc=True, mask_output=False):
        """The interval loop of Wacomm::run over a host window (Wacomm.cpp:200-770).
        Returns (conc [T][s_rho][eta][xi] or None, RunStats)."""
        t = _f64(times)
        arrs = [_f32(a) for a in (zeta, u, v, w, akt)]
        desc = WindowDesc(len(t), _p(t).value, *[_p(a).value for a in arrs])
        src = (Source * max(1, len(sources)))(*sources)
        s_rho, eta, xi = self.shape
        out = np.zeros((len(t), s_rho, eta, xi), np.float32) if conc else None
        st = RunStats()
        self._chk(self.L.wc_run_window(self.h, C.byref(config), C.byref(desc), src, len(sources), PATH[path],
                                       None if out is None else _p(out), int(mask_output), C.byref(st)))
        return out, st

    # -- batched ensemble ---------------------------------------------------
    def ens_set(self, particles: np.ndarray, scenario, sigma):
        """An ensemble population on the current grid: particles (PARTICLE_DTYPE),
        scenario (uint8 per particle), sigma (per scenario)."""
        p = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        sc = np.ascontiguousarray(scenario, dtype=np.uint8)
        sg = _f64(sigma)
        assert len(sc) == len(p)
        self._chk(self.L.wc_ens_set(self.h, _p(p), _p(sc), len(p), _p(sg), len(sg)))
        self.ens_shape = (len(sg),) + self.shape[1:]
        self.ens_n = len(p)

    def ens_release(self, sources, sigma, seed=5489):
        """Device-generated ensemble population from EnsSource records."""
        arr = (EnsSource * max(1, len(sources)))(*sources)
        sg = _f64(sigma)
        self._chk(self.L.wc_ens_release(self.h, arr, len(sources), _p(sg), len(sg), seed))
        self.ens_shape = (len(sg),) + self.shape[1:]
        self.ens_n = int(sum(s.count for s in sources))

    def ens_set_launch(self, block=256, ppt=1, occ=0):
        self._chk(self.L.wc_ens_set_launch(self.h, block, ppt, occ))

    def ens_set_sort(self, tile=8, period=1):
        self._chk(self.L.wc_ens_set_sort(self.h, tile, period))

    def ens_set_lds(self, chunk: int, halo: int = 2):
        """Forcing staged in LDS per workgroup, restaged every `chunk`
        substeps (0: off); the results do not change."""
        self._chk(self.L.wc_ens_set_lds(self.h, int(chunk), int(halo)))

    def ens_get_lds(self):
        c, h = C.c_int(), C.c_int()
        self._chk(self.L.wc_ens_get_lds(self.h, C.byref(c), C.byref(h)))
        return c.value, h.value

    def ens_get_sort(self):
        t, p = C.c_int(), C.c_int()
        self._chk(self.L.wc_ens_get_sort(self.h, C.byref(t), C.byref(p)))
        return t.value, p.value

    def ens_get_launch(self):
        b, p, o = C.c_int(), C.c_int(), C.c_int()
        self._chk(self.L.wc_ens_get_launch(self.h, C.byref(b), C.byref(p), C.byref(o)))
        return b.value, p.value, o.value

    def ens_autotune(self, config: ConfigData, start, end, slot0, slot1):
        ms = C.c_double()
        self._chk(self.L.wc_ens_autotune(self.h, C.byref(config), float(start), float(end), slot0, slot1,
                                         C.byref(ms)))
        return ms.value

    def ens_run(self, config: ConfigData, times, slots, use_graph=True, surface=True) -> EnsStats:
        t = _f64(times)
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        assert len(t) == len(sl) and len(t) >= 1
        st = EnsStats()
        self._chk(self.L.wc_ens_run(self.h, C.byref(config), _p(t), _p(sl), len(t) - 1, int(use_graph),
                                    int(surface), C.byref(st)))
        return st

    def ens_surface(self):
        out = np.zeros(self.ens_shape, np.uint32)
        self._chk(self.L.wc_ens_surface(self.h, _p(out)))
        return out

    def ens_get(self):
        out = np.zeros(self.ens_n, dtype=PARTICLE_DTYPE)
        sc = np.zeros(self.ens_n, np.uint8)
        self._chk(self.L.wc_ens_get(self.h, _p(out), _p(sc), self.ens_n))
        return out, sc

    # -- measurement --------------------------------------------------------
    def set_timing(self, on: bool = True):
        self._chk(self.L.wc_set_timing(self.h, int(on)))

    def kernel_time(self, kernel: int = 0):
        ms, n = C.c_double(), C.c_long()
        self._chk(self.L.wc_kernel_time(self.h, kernel, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def reset_kernel_time(self):
        self._chk(self.L.wc_reset_kernel_time(self.h))

    def step_count(self) -> int:
        v = C.c_uint64()
        self._chk(self.L.wc_step_count(self.h, C.byref(v)))
        return v.value

    def reset_step_count(self):
        self._chk(self.L.wc_reset_step_count(self.h))


def cell_metric_table(lonRad, latRad):
    """Host haversine table of the library (Particle.cpp:544-574)."""
    L = load_library()
    lon, lat = _f64(lonRad), _f64(latRad)
    eta, xi = lon.shape
    dx, dy = np.zeros_like(lon), np.zeros_like(lon)
    L.wc_cell_metric_table(eta, xi, _p(lon), _p(lat), _p(dx), _p(dy))
    return dx, dy


def philox4x32_10(ctr, key):
    L = load_library()
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    L.wc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


class UNet:
    """The concentration-reconstruction network (include/wacomm_hip.h,
    section 4) on one device: images are [n][h][w] float32 arrays, or
    [n][channels][h][w] with `channels` input channels (channel 0 feeds the
    residual); outputs are [n][h][w]."""

    def __init__(self, device: int = 0, seed: int = 1, channels: int = 1):
        self.L = load_library()
        h = vp()
        rc = self.L.wc_unet_create_ex(device, seed, int(channels), C.byref(h))
        if rc:
            raise WcError(rc, f"wc_unet_create_ex(device={device}, channels={channels})")
        self.h = h
        self.channels = int(channels)
        self.n_params = self.L.wc_unet_param_count(h)

    def _chk(self, rc):
        if rc:
            raise WcError(rc, self.L.wc_unet_last_error(self.h).decode(errors="replace"))

    def _shape(self, xx):
        """(n, h, w) of an input batch, checked against the channel count."""
        if self.channels == 1 and xx.ndim == 3:
            return xx.shape
        if xx.ndim != 4 or xx.shape[1] != self.channels:
            raise ValueError(f"input of shape {xx.shape} for a {self.channels}-channel network")
        return xx.shape[0], xx.shape[2], xx.shape[3]

    def close(self):
        if self.h:
            self.L.wc_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph(self, on: bool):
        """Replay the forward and backward passes of gradients() and
        train_step() from a captured graph (the default) or launch them one
        by one; the results are identical."""
        self._chk(self.L.wc_unet_set_graph(self.h, 1 if on else 0))

    def get_params(self):
        out = np.zeros(self.n_params, np.float32)
        self._chk(self.L.wc_unet_get_params(self.h, _p(out)))
        return out

    def set_params(self, params):
        p = _f32(params)
        assert p.size == self.n_params
        self._chk(self.L.wc_unet_set_params(self.h, _p(p)))

    def forward(self, x):
        """(y, device ms of the forward pass)."""
        xx = _f32(x)
        n, hh, ww = self._shape(xx)
        y = np.zeros((n, hh, ww), np.float32)
        ms = C.c_double()
        self._chk(self.L.wc_unet_forward(self.h, _p(xx), n, hh, ww, _p(y), C.byref(ms)))
        return y, ms.value

    def _weight(self, weight, shape):
        if weight is None:
            return None
        w = _f32(weight)
        assert w.shape == shape
        return w

    def gradients(self, x, t, weight=None):
        """(loss, gradient) of mean weight (network(x) - t)^2 (weight None: 1)."""
        xx, tt = _f32(x), _f32(t)
        n, hh, ww = self._shape(xx)
        assert tt.shape == (n, hh, ww)
        w = self._weight(weight, tt.shape)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        self._chk(self.L.wc_unet_gradients_weighted(self.h, _p(xx), _p(tt), None if w is None else _p(w), n, hh, ww,
                                                    _p(g), C.byref(loss)))
        return loss.value, g

    def train_step(self, x, t, lr, weight=None):
        xx, tt = _f32(x), _f32(t)
        n, hh, ww = self._shape(xx)
        assert tt.shape == (n, hh, ww)
        w = self._weight(weight, tt.shape)
        loss = C.c_float()
        self._chk(self.L.wc_unet_train_step_weighted(self.h, _p(xx), _p(tt), None if w is None else _p(w), n, hh, ww,
                                                     float(lr), C.byref(loss)))
        return loss.value


def write_concentration_nc(path, ocean_time, s_rho_values, lat_rho, lon_rho, mask_rho, h, zeta=None, conc=None,
                           attributes=None):
    """Wacomm::save's concentration file (Wacomm.cpp:913-1020) in NetCDF
    classic format: zeta [T][eta][xi], conc [T][s_rho][eta][xi] float (None:
    written as _FillValue), grid fields [eta][xi] in degrees."""
    L = load_library()
    t, sr = _f64(ocean_time), _f64(s_rho_values)
    lat, lon, m, hh = (_f64(a) for a in (lat_rho, lon_rho, mask_rho, h))
    eta, xi = lat.shape
    z = None if zeta is None else _f32(zeta)
    c = None if conc is None else _f32(conc)
    if z is not None:
        assert z.shape == (len(t), eta, xi)
    if c is not None:
        assert c.shape == (len(t), len(sr), eta, xi)
    names = []
    for k, v in (attributes or {}).items():
        names += [str(k).encode(), str(v).encode()]
    arr = (C.c_char_p * (len(names) + 2))(*names, None, None)
    f = ConcFile(len(t), len(sr), eta, xi, _p(t).value, _p(sr).value, _p(lat).value, _p(lon).value, _p(m).value,
                 _p(hh).value, None if z is None else _p(z).value, None if c is None else _p(c).value,
                 C.cast(arr, C.c_void_p).value)
    rc = L.wc_write_concentration_nc(str(path).encode(), C.byref(f))
    if rc:
        raise WcError(rc, f"writing {path}")


def write_restart_nc(path, records: dict, particle_time: float, random_seed: int, backward: bool = False,
                     ocean_model: str = "roms", attributes=None, object_type=None, drift_side=None):
    """Particles::saveAsNetCDF's restart file (Particles.cpp:150-264), restart
    version 3, in the NetCDF classic format with 64-bit data (CDF-5).
    records: the RECORD_FIELDS arrays (as particle_records returns them);
    object_type / drift_side default to the records' own, if they hold them."""
    L = load_library()
    n = len(records["id"])
    if object_type is None:
        object_type = records.get("object_type")
    if drift_side is None:
        drift_side = records.get("drift_side")
    ids = np.ascontiguousarray(records["id"], dtype=np.uint64)
    cols = {f: _f64(records[f]) for f in RECORD_FIELDS if f != "id"}
    for a in cols.values():
        assert a.shape == (n,)
    ot = None if object_type is None else np.ascontiguousarray(object_type, dtype=np.uint32)
    ds = None if drift_side is None else np.ascontiguousarray(drift_side, dtype=np.int32)
    names = []
    for k, v in (attributes or {}).items():
        names += [str(k).encode(), str(v).encode()]
    arr = (C.c_char_p * (len(names) + 2))(*names, None, None)
    pt = lambda a: None if a is None or len(a) == 0 else _p(a).value
    f = RestartFile(n, float(particle_time), int(random_seed), 1 if backward else 0, ocean_model.encode(),
                    C.cast(arr, C.c_void_p).value, pt(ids), pt(ot), pt(ds),
                    *[pt(cols[c]) for c in ("lat", "lon", "depth", "k", "j", "i", "health", "age", "time")])
    rc = L.wc_write_restart_nc(str(path).encode(), C.byref(f))
    if rc:
        raise WcError(rc, f"writing {path}")


def read_restart_nc(path, config: ConfigData, ocean_model: str = "roms"):
    """Particles::loadFromNetCDF (Particles.cpp:373-480): (particles, checkpoint
    time or None) from a NetCDF classic restart file, after the reference's
    metadata checks against config and ocean_model."""
    L = load_library()
    n = C.c_size_t()
    cp = C.c_double(float("nan"))
    err = C.create_string_buffer(512)
    rc = L.wc_read_restart_nc(str(path).encode(), C.byref(config), ocean_model.encode(), None, 0, C.byref(n),
                              C.byref(cp), err, 512)
    if rc:
        raise WcError(rc, err.value.decode(errors="replace"))
    out = np.zeros(n.value, dtype=PARTICLE_DTYPE)
    rc = L.wc_read_restart_nc(str(path).encode(), C.byref(config), ocean_model.encode(),
                              _p(out).value if n.value else None, n.value, C.byref(n), C.byref(cp), err, 512)
    if rc:
        raise WcError(rc, err.value.decode(errors="replace"))
    return out, (None if np.isnan(cp.value) else cp.value)

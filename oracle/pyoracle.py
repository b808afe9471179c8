"""ctypes binding of the CPU oracle (oracle/wc_oracle.c) and of oracle/_ref.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker.  The product
(wacommplusplus_amd) never imports it.

Structures mirror the reference's config_data (Config.hpp:36-72) and
particle_data (Particle.hpp:31-41) field for field.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class CalibrationRegion(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "west", "south", "east", "north", "windErrorStdDev", "windErrorComponentCorrelation",
        "windErrorSpatialScale", "windErrorTemporalScale")]


class Config(C.Structure):
    """config_data (Config.hpp:36-72)."""
    _fields_ = [
        ("random", C.c_bool), ("randomSources", C.c_bool),
        ("deltat", C.c_double), ("dti", C.c_double), ("survprob", C.c_double), ("tau0", C.c_double),
        ("sv", C.c_double), ("crid", C.c_double), ("sigma", C.c_double), ("shoreLimit", C.c_double),
        ("upperClosure", C.c_int), ("lowerClosure", C.c_int), ("horizontalClosure", C.c_int),
        ("randomSeed", C.c_uint64), ("trackingDirection", C.c_int), ("backwardDiffusion", C.c_int),
        ("integrator", C.c_int), ("restartCheckpoint", C.c_double), ("driftModel", C.c_int),
        ("leewayCoefficientEnsemble", C.c_bool), ("leewayResidualCorrelation", C.c_double),
        ("leewayRandomSide", C.c_bool), ("leewayRightSideProbability", C.c_double),
        ("leewayJibeProbabilityHourly", C.c_double), ("windErrorStdDev", C.c_double),
        ("windErrorComponentCorrelation", C.c_double), ("windErrorSpatialScale", C.c_double),
        ("windErrorTemporalScale", C.c_double), ("observationalCalibrationRegionCount", C.c_uint32),
        ("observationalCalibrationRegions", CalibrationRegion * 16),
        ("driftObjectType", C.c_uint16), ("driftSide", C.c_int8), ("hasWind", C.c_bool),
        ("windU10", C.c_double), ("windV10", C.c_double),
    ]


CONSTRAINT, KILL, REFLECTION = 1, 2, 3     # Config.hpp:201-203
EULER, MIDPOINT = 1, 2                     # Config.hpp:206-207


def default_config(**kw) -> Config:
    """Config::setDefault values that the move reads (Config.cpp:46-202)."""
    c = Config()
    c.random = True
    c.randomSources = True
    c.deltat = 3600.0
    c.dti = 30.0
    c.survprob = 1.0e-4
    c.tau0 = 86400.0
    c.sv = 0.0
    c.crid = 1.0
    c.sigma = 3.46
    c.shoreLimit = 0.25
    c.upperClosure = CONSTRAINT
    c.lowerClosure = KILL
    c.horizontalClosure = REFLECTION
    c.randomSeed = 5489
    c.trackingDirection = 1
    c.backwardDiffusion = 0
    c.integrator = MIDPOINT
    c.restartCheckpoint = math.nan
    c.driftModel = 0
    for k, v in kw.items():
        setattr(c, k, v)
    return c


# numpy view of particle_data (64 bytes, Particle.hpp:31-41)
PARTICLE_DTYPE = np.dtype({
    "names": ["id", "k", "j", "i", "health", "age", "time", "driftObjectType", "driftSide"],
    "formats": [np.uint64, np.float64, np.float64, np.float64, np.float64, np.float64, np.float64,
                np.uint16, np.int8],
    "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 58],
    "itemsize": 64,
})


def make_particles(n, k=0.0, j=0.0, i=0.0, time=0.0, health=1.0, age=0.0, first_id=0):
    p = np.zeros(n, dtype=PARTICLE_DTYPE)
    p["id"] = np.arange(first_id, first_id + n, dtype=np.uint64)
    p["k"], p["j"], p["i"] = k, j, i
    p["health"], p["age"], p["time"] = health, age, time
    return p


class _Grid(C.Structure):
    _fields_ = [("s_w", C.c_int), ("s_rho", C.c_int), ("eta", C.c_int), ("xi", C.c_int),
                ("lonRad", C.c_void_p), ("latRad", C.c_void_p), ("h", C.c_void_p),
                ("depthIntervals", C.c_void_p)]


class _Record(C.Structure):
    _fields_ = [("zeta", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p),
                ("akt", C.c_void_p)]


class _RomsGrid(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("s_rho", "eta", "xi", "eta_u", "xi_u", "eta_v", "xi_v")] + \
               [(n, C.c_void_p) for n in ("mask_rho", "mask_u", "mask_v", "angle")] + \
               [("earth_relative", C.c_int)] + \
               [(n, C.c_void_p) for n in ("missing_u", "missing_v", "missing_angle")] + \
               [(n, C.c_int) for n in ("n_missing_u", "n_missing_v", "n_missing_angle")]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _records(particles):
    """particle_data records as the C side reads them: 64-byte, contiguous
    (numpy can pack a structured array, e.g. in concatenate)."""
    if particles.dtype.itemsize != 64 or not particles.flags["C_CONTIGUOUS"]:
        raise ValueError(f"particles must be contiguous 64-byte particle_data records, got {particles.dtype.itemsize}-"
                         f"byte records (use .astype(PARTICLE_DTYPE))")
    return _ptr(particles)


def _load(path):
    lib = C.CDLL(path)
    lib.wco_mix.restype = C.c_uint64
    lib.wco_mix.argtypes = [C.c_uint64]
    for name in ("wco_uniform", "wco_normal"):
        f = getattr(lib, name)
        f.restype = C.c_double
        f.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_uint64, C.c_uint64]
    lib.wco_jibe_step_probability.restype = C.c_double
    lib.wco_jibe_step_probability.argtypes = [C.c_double, C.c_double]
    lib.wco_sample_drift_side.restype = C.c_int
    lib.wco_sample_drift_side.argtypes = [C.c_uint64, C.c_uint64, C.c_double]
    lib.wco_restart_elapsed.restype = C.c_double
    lib.wco_restart_elapsed.argtypes = [C.c_double] * 3
    lib.wco_reflect_cell.restype = C.c_double
    lib.wco_reflect_cell.argtypes = [C.c_double, C.c_int, C.c_double, C.c_int]
    lib.wco_reflect_domain.restype = C.c_double
    lib.wco_reflect_domain.argtypes = [C.c_double, C.c_double]
    lib.wco_move.argtypes = [C.POINTER(Config), C.POINTER(_Grid), C.c_double, C.c_double,
                             C.POINTER(_Record), C.POINTER(_Record), C.c_void_p]
    lib.wco_move_all.argtypes = [C.POINTER(Config), C.POINTER(_Grid), C.c_double, C.c_double,
                                 C.POINTER(_Record), C.POINTER(_Record), C.c_void_p, C.c_long, C.c_int]
    lib.wco_grid_count.argtypes = [C.c_void_p, C.c_long, C.c_int, C.c_int, C.c_int, C.c_void_p]
    lib.wco_emission_times.restype = C.c_long
    lib.wco_emission_times.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_void_p, C.c_long]
    lib.wco_cell_metric.argtypes = [C.POINTER(_Grid), C.c_int, C.c_int, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
    lib.wco_roms_normalize.restype = C.c_int
    lib.wco_roms_normalize.argtypes = [C.POINTER(_RomsGrid)] + [C.c_void_p] * 10
    lib.wco_roms_static.argtypes = [C.c_int] * 3 + [C.c_void_p] * 9
    lib.wco_kji2deplatlon.argtypes = [C.c_double] * 3 + [C.c_int] * 3 + [C.c_void_p] * 7
    lib.wco_particle_records.restype = C.c_long
    lib.wco_particle_records.argtypes = [C.c_void_p, C.c_long] + [C.c_int] * 3 + [C.c_void_p] * 8
    return lib


_LIBS = {}


def lib(plant: int = 0):
    """The oracle library; plant>0 selects a deliberately wrong build."""
    if plant not in _LIBS:
        path = os.path.join(HERE, "liboracle.so") if plant == 0 else \
            os.path.join(HERE, "_plant", f"liboracle_plant{plant}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        _LIBS[plant] = _load(path)
    return _LIBS[plant]


def ref_lib():
    """oracle/_ref/libwcref.so (reference headers compiled in place) or None."""
    path = os.path.join(HERE, "_ref", "libwcref.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.ref_mix.restype = C.c_uint64
    L.ref_mix.argtypes = [C.c_uint64]
    for name in ("ref_uniform", "ref_normal"):
        f = getattr(L, name)
        f.restype = C.c_double
        f.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_uint64, C.c_uint64]
    L.ref_restart_elapsed.restype = C.c_double
    L.ref_restart_elapsed.argtypes = [C.c_double] * 3
    L.ref_reflect_cell.restype = C.c_double
    L.ref_reflect_cell.argtypes = [C.c_double, C.c_int, C.c_double, C.c_int]
    L.ref_reflect_domain.restype = C.c_double
    L.ref_reflect_domain.argtypes = [C.c_double, C.c_double]
    L.ref_scheme.restype = C.c_char_p
    L.ref_emission_times.restype = C.c_long
    L.ref_emission_times.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_void_p, C.c_long]
    L.ref_source_position_offset.restype = C.c_double
    L.ref_source_position_offset.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    if hasattr(L, "ref_jibe_step_probability"):
        L.ref_jibe_step_probability.restype = C.c_double
        L.ref_jibe_step_probability.argtypes = [C.c_double, C.c_double]
        L.ref_sample_drift_side.restype = C.c_int
        L.ref_sample_drift_side.argtypes = [C.c_uint64, C.c_uint64, C.c_double]
        L.ref_jibe_component.restype = C.c_uint64
    if hasattr(L, "ref_partition_count"):
        for name in ("ref_partition_count", "ref_partition_offset"):
            f = getattr(L, name)
            f.restype = C.c_uint64
            f.argtypes = [C.c_uint64] * 3
    return L


def partition(total, workers, worker):
    """(offset, count) of `worker`'s contiguous particle range when `total`
    particles are split over `workers` ranks: every rank gets total // workers
    and rank 0 also the remainder (NumericalHelpers::partitionCount /
    partitionOffset, NumericalHelpers.hpp:76-82; MpiDecompositionTest.cpp:16-20
    scatters by exactly these ranges)."""
    base = total // workers
    count = base + (total % workers if worker == 0 else 0)
    offset = 0 if worker == 0 else base + total % workers + (worker - 1) * base
    return offset, count


class Window:
    """A forcing window held on the host: static grid + T records (FP32 fields).

    Layout (NumericalHelpers.hpp:16-18, cuda/kernel.cu:63-69):
      zeta[T][eta][xi], u,v[T][s_rho][eta][xi], w,akt[T][s_w][eta][xi];
      h, lonRad, latRad[eta][xi]; depthIntervals[s_w-1] (entry kI+s_w-2).
    """

    def __init__(self, ocean_time, lonRad, latRad, h, depthIntervals, zeta, u, v, w, akt, mask=None):
        self.ocean_time = np.ascontiguousarray(ocean_time, dtype=np.float64)
        self.lonRad = np.ascontiguousarray(lonRad, dtype=np.float64)
        self.latRad = np.ascontiguousarray(latRad, dtype=np.float64)
        self.h = np.ascontiguousarray(h, dtype=np.float64)
        self.depthIntervals = np.ascontiguousarray(depthIntervals, dtype=np.float64)
        self.zeta = np.ascontiguousarray(zeta, dtype=np.float32)
        self.u = np.ascontiguousarray(u, dtype=np.float32)
        self.v = np.ascontiguousarray(v, dtype=np.float32)
        self.w = np.ascontiguousarray(w, dtype=np.float32)
        self.akt = np.ascontiguousarray(akt, dtype=np.float32)
        self.mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.float64)
        self.T = self.zeta.shape[0]
        self.eta, self.xi = self.h.shape
        self.s_rho = self.u.shape[1]
        self.s_w = self.w.shape[1]
        assert self.depthIntervals.shape[0] >= self.s_w - 1
        assert self.u.shape == (self.T, self.s_rho, self.eta, self.xi) == self.v.shape
        assert self.w.shape == (self.T, self.s_w, self.eta, self.xi) == self.akt.shape
        self._grid = _Grid(self.s_w, self.s_rho, self.eta, self.xi, _ptr(self.lonRad).value,
                           _ptr(self.latRad).value, _ptr(self.h).value, _ptr(self.depthIntervals).value)

    def record(self, t):
        return _Record(_ptr(self.zeta[t]).value, _ptr(self.u[t]).value, _ptr(self.v[t]).value,
                       _ptr(self.w[t]).value, _ptr(self.akt[t]).value)

    def interval(self, ocean_time_idx, direction, deltat=3600.0):
        """(start, end, record index pair) as Particle.cpp:110-115,183-184."""
        nxt = ocean_time_idx + direction
        start = float(self.ocean_time[ocean_time_idx])
        end = float(self.ocean_time[nxt]) if 0 <= nxt < self.T else start + direction * deltat
        if nxt < 0 or nxt >= self.T:
            nxt = ocean_time_idx
        return start, end, ocean_time_idx, nxt

    def cell_metric(self, j, i, plant=0):
        dx, dy = C.c_double(), C.c_double()
        lib(plant).wco_cell_metric(C.byref(self._grid), j, i, C.byref(dx), C.byref(dy))
        return dx.value, dy.value


def move(window: Window, cfg: Config, ocean_time_idx: int, particles: np.ndarray, nthreads: int = 0,
         plant: int = 0):
    """One forcing interval of Particle::move over every particle, in place."""
    assert particles.dtype == PARTICLE_DTYPE and particles.flags.c_contiguous
    s, e, t0, t1 = window.interval(ocean_time_idx, cfg.trackingDirection, cfg.deltat)
    r0, r1 = window.record(t0), window.record(t1)
    lib(plant).wco_move_all(C.byref(cfg), C.byref(window._grid), s, e, C.byref(r0), C.byref(r1),
                            _records(particles), len(particles), nthreads)
    return particles


def grid_count(particles: np.ndarray, s_rho: int, eta: int, xi: int, plant: int = 0) -> np.ndarray:
    counts = np.zeros((s_rho, eta, xi), dtype=np.uint32)
    lib(plant).wco_grid_count(_records(particles), len(particles), s_rho, eta, xi, _ptr(counts))
    return counts


def emission_times(mode, value, source_start, source_end, interval_start, interval_end, cap=1 << 20,
                   use_ref=False):
    buf = np.zeros(cap, dtype=np.float64)
    L = ref_lib() if use_ref else lib()
    f = L.ref_emission_times if use_ref else L.wco_emission_times
    n = f(mode, value, source_start, source_end, interval_start, interval_end, _ptr(buf), cap)
    if n < 0:
        return n
    return buf[:n].copy()


class RomsGrid:
    """The static ROMS grid of one file: rho/u/v masks, the grid angle and
    the missing-value attributes (ROMSAdapter.cpp:63-166)."""

    def __init__(self, s_rho, mask_rho, mask_u, mask_v, angle=None, earth_relative=False,
                 missing_u=(), missing_v=(), missing_angle=()):
        self.mask_rho = np.ascontiguousarray(mask_rho, dtype=np.float64)
        self.mask_u = np.ascontiguousarray(mask_u, dtype=np.float64)
        self.mask_v = np.ascontiguousarray(mask_v, dtype=np.float64)
        eta, xi = self.mask_rho.shape
        self.angle = np.ascontiguousarray(np.zeros((eta, xi)) if angle is None else angle, dtype=np.float64)
        self.missing = [np.ascontiguousarray(np.r_[m, 0.0][: max(1, len(m))], dtype=np.float64)
                        for m in (missing_u, missing_v, missing_angle)]
        self.counts = [len(m) for m in (missing_u, missing_v, missing_angle)]
        self.s_rho, self.eta, self.xi = s_rho, eta, xi
        self._g = _RomsGrid(s_rho, eta, xi, *self.mask_u.shape, *self.mask_v.shape,
                            _ptr(self.mask_rho).value, _ptr(self.mask_u).value, _ptr(self.mask_v).value,
                            _ptr(self.angle).value, int(bool(earth_relative)),
                            *[_ptr(m).value for m in self.missing], *self.counts)


def roms_normalize(g: RomsGrid, zeta, u, v, w, akt):
    """One ROMS record onto the rho grid: (rc, zeta, U, V, W, AKT)."""
    ins = [np.ascontiguousarray(a, dtype=np.float32) for a in (zeta, u, v, w, akt)]
    assert ins[1].shape == (g.s_rho,) + g.mask_u.shape and ins[2].shape == (g.s_rho,) + g.mask_v.shape
    assert ins[0].shape == (g.eta, g.xi) and ins[3].shape == ins[4].shape == (g.s_rho + 1, g.eta, g.xi)
    outs = [np.zeros((g.eta, g.xi), np.float32), np.zeros((g.s_rho, g.eta, g.xi), np.float32),
            np.zeros((g.s_rho, g.eta, g.xi), np.float32), np.zeros((g.s_rho + 1, g.eta, g.xi), np.float32),
            np.zeros((g.s_rho + 1, g.eta, g.xi), np.float32)]
    rc = lib().wco_roms_normalize(C.byref(g._g), *[_ptr(a) for a in ins], *[_ptr(a) for a in outs])
    return (rc, *outs)


def roms_static(s_w_values, lon, lat, h, mask_rho):
    """(lonRad, latRad, h, depthIntervals) of ROMSAdapter.cpp:232-275."""
    sw = np.ascontiguousarray(s_w_values, dtype=np.float64)
    lon, lat, h, m = (np.ascontiguousarray(a, dtype=np.float64) for a in (lon, lat, h, mask_rho))
    eta, xi = m.shape
    lr, la, ho = np.zeros((eta, xi)), np.zeros((eta, xi)), np.zeros((eta, xi))
    di = np.zeros(len(sw) - 1)
    lib().wco_roms_static(len(sw) - 1, eta, xi, _ptr(sw), _ptr(lon), _ptr(lat), _ptr(h), _ptr(m), _ptr(lr),
                          _ptr(la), _ptr(ho), _ptr(di))
    return lr, la, ho, di


def kji2deplatlon(k, j, i, s_w_values, lat, lon, h):
    """(depth, lat, lon) of OceanModelAdapter::kji2deplatlon (OceanModelAdapter.cpp:271-321)."""
    sw = np.ascontiguousarray(s_w_values, dtype=np.float64)
    lat, lon, h = (np.ascontiguousarray(a, dtype=np.float64) for a in (lat, lon, h))
    eta, xi = h.shape
    d, la, lo = C.c_double(), C.c_double(), C.c_double()
    lib().wco_kji2deplatlon(float(k), float(j), float(i), len(sw), eta, xi, _ptr(sw), _ptr(lat), _ptr(lon),
                            _ptr(h), C.byref(d), C.byref(la), C.byref(lo))
    return d.value, la.value, lo.value


def particle_records(particles: np.ndarray, s_w_values, lat, lon, h):
    """Particles::netcdfRecords (Particles.cpp:115-147): (src indices, depth, lat, lon)."""
    particles = np.ascontiguousarray(particles)
    sw = np.ascontiguousarray(s_w_values, dtype=np.float64)
    lat, lon, h = (np.ascontiguousarray(a, dtype=np.float64) for a in (lat, lon, h))
    eta, xi = h.shape
    n = len(particles)
    src = np.zeros(max(n, 1), np.int64)
    dep, la, lo = (np.zeros(max(n, 1)) for _ in range(3))
    r = lib().wco_particle_records(_records(particles), n, len(sw), eta, xi, _ptr(sw), _ptr(lat), _ptr(lon), _ptr(h),
                                   _ptr(src), _ptr(dep), _ptr(la), _ptr(lo))
    return src[:r], dep[:r], la[:r], lo[:r]

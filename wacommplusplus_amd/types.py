"""Host-side data types of the C ABI (include/wacomm_hip.h).

ConfigData is field-for-field the reference's config_data (Config.hpp:36-72);
PARTICLE_DTYPE is particle_data (Particle.hpp:31-41, 64 bytes) as a numpy
structured dtype, which is also the layout of a std::vector<particle_data>.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

CONSTRAINT, KILL, REFLECTION = 1, 2, 3     # Config.hpp:201-203
EULER, MIDPOINT = 1, 2                     # Config.hpp:206-207
FORWARD, BACKWARD = 1, -1                  # Config.hpp:204-205

STORE = {"f32": 0, "f16": 1, "bf16": 2, "bf16_tile": 3, "f16_tile": 4, "i16_tile": 5}
# closure-event bits of wc_get_events
EVENT = {"shore": 1, "domain": 2, "fallback": 4, "vertical": 8, "shallow": 16, "coast": 32}
PATH = {"parity": 0, "fast": 1}


class CalibrationRegion(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "west", "south", "east", "north", "windErrorStdDev", "windErrorComponentCorrelation",
        "windErrorSpatialScale", "windErrorTemporalScale")]


class ConfigData(C.Structure):
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

    def copy(self, **kw) -> "ConfigData":
        c = ConfigData()
        C.pointer(c)[0] = self
        for k, v in kw.items():
            setattr(c, k, v)
        return c


def default_config(**kw) -> ConfigData:
    """The Config::setDefault values the move reads (Config.cpp:46-202)."""
    c = ConfigData()
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
    c.trackingDirection = FORWARD
    c.backwardDiffusion = 0
    c.integrator = MIDPOINT
    c.restartCheckpoint = math.nan
    c.driftModel = 0
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def sigma_from_kh(kh: float, dti: float) -> float:
    """Horizontal displacement scale whose per-step variance sigma^2 gives an
    effective diffusivity K_h = sigma^2 / (2 dti) (docs/model.md:204-210)."""
    return math.sqrt(2.0 * kh * dti)


PARTICLE_DTYPE = np.dtype({
    "names": ["id", "k", "j", "i", "health", "age", "time", "driftObjectType", "driftSide"],
    "formats": [np.uint64, np.float64, np.float64, np.float64, np.float64, np.float64, np.float64,
                np.uint16, np.int8],
    "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 58],
    "itemsize": 64,
})


def make_particles(n, k=0.0, j=0.0, i=0.0, time=0.0, health=1.0, age=0.0, first_id=0) -> np.ndarray:
    """Particles as Particle(id, k, j, i, time) constructs them (Particle.cpp:30-44)."""
    p = np.zeros(n, dtype=PARTICLE_DTYPE)
    p["id"] = np.arange(first_id, first_id + n, dtype=np.uint64)
    p["k"], p["j"], p["i"] = k, j, i
    p["health"], p["age"], p["time"] = health, age, time
    return p


# (cout, c1, c2) of the reconstruction network's 3x3 conv layers, in parameter
# order (include/wacomm_hip.h, section 4)
UNET_SPEC = [(32, 32, 0), (32, 32, 0), (64, 32, 0), (64, 64, 0), (128, 64, 0), (128, 128, 0), (256, 128, 0),
             (256, 256, 0), (128, 128, 256), (128, 128, 0), (64, 64, 128), (64, 64, 0), (32, 32, 64), (32, 32, 0)]

"""wacommplusplus_amd -- MI355X (gfx950) particle move and concentration
gridding for WaComM++, behind the reference's Particles/Wacomm roles.

The compute path is libwacomm_hip.so (hand-written HIP kernels, C ABI in
include/wacomm_hip.h); this package is its host-side binding plus synthetic
workload generators.  There is no CPU fallback: importing the binding fails
loudly when the library has not been built.
"""
from .types import (BACKWARD, CONSTRAINT, EULER, EVENT, FORWARD, KILL, MIDPOINT, PARTICLE_DTYPE, REFLECTION, STORE,
                    ConfigData, default_config, make_particles, sigma_from_kh)
from ._lib import (EXPORTS, FORCING_INTERVAL_BATCH, LIB_PATH, ROMS_MAX_MISSING, SINGLE_PULSE, UNIFORM_RATE,
                   WC_ERR_HIP, WC_ERR_INVALID, WC_ERR_NO_DEVICE, WC_ERR_NOMEM, WC_ERR_UNSUPPORTED, WC_OK, Context, EnsSource, EnsStats,
                   RunStats, Source, UNet, WcError, cell_metric_table, load_library, philox4x32_10,
                   write_concentration_nc, WC_ERR_IO, RECORD_FIELDS, read_restart_nc, write_restart_nc)

__all__ = [
    "BACKWARD", "CONSTRAINT", "EULER", "FORWARD", "KILL", "MIDPOINT", "PARTICLE_DTYPE", "REFLECTION",
    "ConfigData", "default_config", "make_particles", "sigma_from_kh", "EXPORTS", "LIB_PATH", "Context",
    "WcError", "cell_metric_table", "load_library", "philox4x32_10", "Source", "RunStats", "UNIFORM_RATE",
    "FORCING_INTERVAL_BATCH", "SINGLE_PULSE", "ROMS_MAX_MISSING", "WC_OK", "WC_ERR_HIP", "WC_ERR_INVALID",
    "WC_ERR_UNSUPPORTED", "WC_ERR_NOMEM", "WC_ERR_NO_DEVICE", "EnsStats", "EnsSource", "UNet", "write_concentration_nc", "WC_ERR_IO",
    "RECORD_FIELDS", "read_restart_nc", "write_restart_nc", "EVENT", "STORE",
]

"""Minimal ctypes access to the HIP runtime that libwacomm_hip.so links
(device buffers for the drop-in entry points, which take device pointers)."""
import ctypes as C

import numpy as np

import wacommplusplus_amd as wc

wc.load_library()                          # loads libamdhip64.so.7 first: the same runtime is reused
_hip = C.CDLL("libamdhip64.so.7")
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_hip.hipDeviceSynchronize.argtypes = []
H2D, D2H = 1, 2


class DeviceBuffer:
    def __init__(self, host: np.ndarray):
        # particle records always travel as the 64-byte particle_data layout
        # (numpy may pack a structured array, e.g. the result of concatenate)
        if host.dtype.names is not None and host.dtype.names == wc.PARTICLE_DTYPE.names:
            host = host.astype(wc.PARTICLE_DTYPE)
        host = np.ascontiguousarray(host)
        self.dtype = host.dtype
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert _hip.hipMalloc(C.byref(self.ptr), max(1, self.nbytes)) == 0
        assert _hip.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p), self.nbytes, H2D) == 0

    def get(self, like: np.ndarray) -> np.ndarray:
        out = np.empty(like.shape, dtype=self.dtype)
        assert out.nbytes == self.nbytes
        assert _hip.hipDeviceSynchronize() == 0
        assert _hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = None


def synchronize():
    assert _hip.hipDeviceSynchronize() == 0


def run_dropin(W, cfg, particles, n_intervals, first=None):
    """The interval loop of Wacomm::run through wc_hip_move_particles, the
    cudaMoveParticle replacement (cuda/kernel.h:7): the whole window in the
    reference's device layout (zeta [T][eta][xi], u/v [T][s_rho][eta][xi],
    w/akt [T][s_w][eta][xi], FP32; times, lon/lat (radians), h, mask FP64),
    config_data and particle_data in device memory, one call per interval
    with timeIndex = the interval's start record."""
    from runner import intervals
    zeta, u, v, w, akt = W.stacked()
    bufs = {k: DeviceBuffer(np.ascontiguousarray(a)) for k, a in dict(
        times=np.asarray(W.times, np.float64), mask=np.asarray(W.mask, np.float64),
        lon=np.asarray(W.lonRad, np.float64), lat=np.asarray(W.latRad, np.float64),
        depth=np.asarray(W.depthIntervals, np.float64), h=np.asarray(W.h, np.float64),
        zeta=zeta.astype(np.float32), u=u.astype(np.float32), v=v.astype(np.float32), w=w.astype(np.float32),
        akt=akt.astype(np.float32)).items()}
    p = np.ascontiguousarray(particles.copy())
    pd = DeviceBuffer(p)
    cd = DeviceBuffer(np.frombuffer(bytes(cfg), np.uint8))
    L = wc.load_library()
    try:
        for t, _ in intervals(W, cfg.trackingDirection, n_intervals, first):
            rc = L.wc_hip_move_particles(cd.ptr, pd.ptr, t, W.T, W.s_w, W.s_rho, W.eta, W.xi, bufs["times"].ptr,
                                         bufs["mask"].ptr, bufs["lon"].ptr, bufs["lat"].ptr, bufs["depth"].ptr,
                                         bufs["h"].ptr, bufs["zeta"].ptr, bufs["u"].ptr, bufs["v"].ptr,
                                         bufs["w"].ptr, bufs["akt"].ptr, None, None, None, None, len(p), 1, 1, None)
            assert rc == wc.WC_OK, rc
        return pd.get(p)
    finally:
        for b in list(bufs.values()) + [pd, cd]:
            b.free()

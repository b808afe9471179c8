"""Error convention of the C ABI (the CUDA backend's caller prints and exits,
Wacomm.cpp:38-46; this library never exits): every rejected call returns its
wc_status with a message from wc_last_error, changes nothing, and leaves the
context usable -- a valid run afterwards still matches the oracle bit for bit.
"""
import ctypes as C
import math

import numpy as np
import pytest

import wacommplusplus_amd as wc
from runner import run_device, run_oracle
from scenarios import tiny_window
from test_gpu_parity import as_window

pytestmark = pytest.mark.gpu
INVALID, UNSUPPORTED = wc.WC_ERR_INVALID, wc.WC_ERR_UNSUPPORTED


def expect(ctx, rc, status):
    assert rc == status, (rc, status)
    msg = ctx.L.wc_last_error(ctx.h).decode()
    assert msg, "no message for a rejected call"
    return msg


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def test_rejected_calls_leave_the_context_usable():
    W = as_window(tiny_window())
    L = wc.load_library()
    p = wc.make_particles(64, k=-.5, j=.25, i=.25)
    p["id"] = np.arange(64, dtype=np.uint64)
    cfg = wc.default_config(random=True, sigma=0.1, randomSeed=3)
    with wc.Context(0) as ctx:
        h = ctx.h
        lon, lat, hh, dep = f64(W.lonRad), f64(W.latRad), f64(W.h), f64(W.depthIntervals)
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)                      # noqa: E731
        # before any grid
        expect(ctx, L.wc_move_interval(h, C.byref(cfg), 0.0, 65.0, 0, 1, 0), INVALID)
        expect(ctx, L.wc_grid(h, None, None, 0), INVALID)
        # malformed grids
        expect(ctx, L.wc_set_grid(h, 4, 2, 2, 2, ptr(lon), ptr(lat), ptr(hh), None, ptr(dep), 2, 0), INVALID)
        expect(ctx, L.wc_set_grid(h, 3, 2, 2, 2, ptr(lon), ptr(lat), ptr(hh), None, ptr(dep), 1, 0), INVALID)
        expect(ctx, L.wc_set_grid(h, 3, 2, 2, 2, ptr(lon), ptr(lat), ptr(hh), None, ptr(dep), 2, 7), INVALID)
        expect(ctx, L.wc_set_grid(h, 3, 2, 2, 2, None, ptr(lat), ptr(hh), None, ptr(dep), 2, 0), INVALID)
        ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask)
        ctx.set_particles(p)
        # forcing slots
        rec = [np.ascontiguousarray(a, dtype=np.float32) for a in W.record(0)]
        expect(ctx, L.wc_upload_record(h, 2, *[ptr(a) for a in rec]), INVALID)
        expect(ctx, L.wc_upload_record(h, -1, *[ptr(a) for a in rec]), INVALID)
        ctx.upload_record(0, *W.record(0))
        msg = expect(ctx, L.wc_move_interval(h, C.byref(cfg), 0.0, 65.0, 0, 1, 0), INVALID)   # slot 1 empty
        assert "slot" in msg
        ctx.upload_record(1, *W.record(1))
        # move arguments
        expect(ctx, L.wc_move_interval(h, C.byref(cfg), 0.0, 65.0, 0, 2, 0), INVALID)
        expect(ctx, L.wc_move_interval(h, C.byref(cfg), 0.0, 65.0, 0, 1, 5), INVALID)
        expect(ctx, L.wc_move_interval(h, C.byref(cfg), math.nan, 65.0, 0, 1, 0), INVALID)
        expect(ctx, L.wc_move_interval(h, C.byref(cfg), 0.0, math.inf, 0, 1, 0), INVALID)
        for bad in (dict(trackingDirection=0), dict(dti=0.0), dict(dti=-30.0)):
            expect(ctx, L.wc_move_interval(h, C.byref(cfg.copy(**bad)), 0.0, 65.0, 0, 1, 0), INVALID)
        assert "leeway" in expect(ctx, L.wc_move_interval(h, C.byref(cfg.copy(driftModel=1)), 0.0, 65.0, 0, 1, 0),
                                  UNSUPPORTED)
        # particle transfers
        out = np.zeros(65, dtype=wc.PARTICLE_DTYPE)
        expect(ctx, L.wc_get_particles(h, ptr(out), 65), INVALID)
        expect(ctx, L.wc_set_particles(h, None, 5), INVALID)
        # nothing changed: the population is the one set, and a valid move
        # matches the oracle
        got = ctx.get_particles()
        assert np.array_equal(got, p)
        dev = run_device(ctx, W, cfg, p, 1)
        ref = run_oracle(W, cfg, p, 1)
        for f in ("k", "j", "i"):
            assert np.max(np.abs(dev[f] - ref[f])) <= 1e-9, f
        assert np.array_equal(dev["age"], ref["age"])


def test_dropin_entry_points_reject_bad_arguments():
    """wc_hip_move_particles / wc_hip_concentration with impossible sizes or
    missing buffers return WC_ERR_INVALID before touching any memory."""
    L = wc.load_library()
    # a missing config and a negative count; a frame missing, a negative count, no levels
    assert L.wc_hip_move_particles(None, None, 0, 2, 3, 2, 2, 2, *([None] * 6), *([None] * 9), -1, 1, 1, None) == INVALID
    assert L.wc_hip_concentration(None, 5, 2, 2, 2, None, None) == INVALID
    assert L.wc_hip_concentration(None, 0, 0, 2, 2, None, None) == INVALID
    assert L.wc_hip_concentration(None, 0, 2, 2, 2, None, None) == INVALID      # no frame

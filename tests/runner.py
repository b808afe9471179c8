"""Drivers shared by the parity tests: run the same forcing window and release
through the oracle (CPU) and through libwacomm_hip.so (GPU) interval by
interval, in the order of Wacomm::run (Wacomm.cpp:204-770)."""
from __future__ import annotations

import numpy as np

import pyoracle as po


def oracle_pair(W, t0, t1):
    """Oracle view of the interval from record t0 to record t1 of a synthetic
    ForcingWindow: the two records in chronological order, as the reference
    holds them, and the time index of the interval start t0 (1 when tracking
    backward: Particle::move goes from ocean_time(idx) to
    ocean_time(idx + direction), Particle.cpp:110-115)."""
    lo, hi = min(t0, t1), max(t0, t1)
    r0, r1 = W.record(lo), W.record(hi)
    stack = [np.stack([r0[q], r1[q]]) for q in range(5)]
    times = np.array([W.times[lo], W.times[hi]])
    return po.Window(times, W.lonRad, W.latRad, W.h, W.depthIntervals, *stack), (0 if t1 >= t0 else 1)


def cfg_to_oracle(cfg):
    c = po.Config()
    import ctypes as C
    C.memmove(C.byref(c), C.byref(cfg), C.sizeof(c))
    return c


def intervals(W, direction, n, first=None):
    """(t, next) pairs of the interval loop (Wacomm.cpp:200-206)."""
    t = (0 if direction > 0 else W.T - 1) if first is None else first
    out = []
    for _ in range(n):
        nxt = t + direction
        assert 0 <= nxt < W.T
        out.append((t, nxt))
        t = nxt
    return out


def run_oracle(W, cfg, particles, n_intervals, first=None, nthreads=0, remove_dead=False):
    p = particles.copy()
    ocfg = cfg_to_oracle(cfg)
    for t, nxt in intervals(W, cfg.trackingDirection, n_intervals, first):
        win, idx = oracle_pair(W, t, nxt)
        po.move(win, ocfg, idx, p, nthreads=nthreads)
        if remove_dead:
            p = p[p["health"] > 0]
    return p


def load_grid(ctx, W, n_slots=2, store="f32"):
    ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask, n_slots=n_slots, store=store)


def run_device(ctx, W, cfg, particles, n_intervals, first=None, path="parity", store="f32", remove_dead=False,
               grid_frames=False):
    """Two-slot record ring: record t lives in slot t % 2."""
    load_grid(ctx, W, 2, store)
    ctx.set_particles(particles)
    frames = []
    pairs = intervals(W, cfg.trackingDirection, n_intervals, first)
    ctx.upload_record(pairs[0][0] % 2, *W.record(pairs[0][0]))
    for t, nxt in pairs:
        ctx.upload_record(nxt % 2, *W.record(nxt))
        start, end = float(W.times[t]), float(W.times[nxt])
        ctx.move_interval(cfg, start, end, t % 2, nxt % 2, path=path)
        if remove_dead:
            ctx.remove_dead()
        if grid_frames:
            frames.append(ctx.grid())
    out = ctx.get_particles()
    return (out, frames) if grid_frames else out


def max_cell_error(a, b, fields=("k", "j", "i")):
    return max(float(np.max(np.abs(a[f] - b[f]))) if len(a) else 0.0 for f in fields)

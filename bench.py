"""Benchmark of the particle-move + gridding path on one MI355X.

Workload (the full-size plain-method configuration): 1024(eta) x 768(xi) x 11
sigma-level grid at 150 m with ~30 % land, u/v from 64 random Fourier modes
(5-40 km, rms 0.3 m/s), w from continuity, 25 hourly records; 4,000,000
particles released at t0 from 50 random ocean point sources at 1 m depth;
Kh = 1 m2/s (sigma = sqrt(2 Kh dti)), AKt = 1e-4 m2/s, dti = 30 s, explicit
midpoint (Euler reported beside it), default closures.  Synthetic data
generated in memory.

A step is one forcing interval of the hot path for the whole population:
move (120 substeps) -> remove dead -> grid (counts, then the reference's float
frame with the land mask, on the device), i.e. one pass of the interval loop
of Wacomm::run.  All forcing records are resident in HBM before the timed
region (the headline `value`); the rate with one record uploaded per interval
from host memory is reported beside it.  The headline is the FP64-state
parity path (FP32-stored forcing, as the reference stores it); the mixed-
precision path with BF16 and FP16 storage (plain and as tile-mean residuals)
and FP32 storage is reported alongside.

Rooflines.  The move keeps each particle in registers for a whole interval and
its forcing gathers are served by L2 and the Infinity Cache, so it is bound by
vector instruction issue, not by HBM: `roofline` is the VALU issue rate (wave
instructions per dispatch from the committed rocprofv3 counters, divided by
the launch time measured here) against the chip's 256 CU x 4 SIMD x 2.4 GHz /
4 cycles.  `roofline.hbm` gives the measured DRAM bytes per dispatch (the
same counters) over the launch time against 8 TB/s, and `effective_gbs` the
reference stencil's algorithmic bytes over the launch time (an L2/MALL-served
figure, not an HBM fraction).

    python bench.py [--gpus 1] [--steps 24] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4     # wave64 VALU instructions per second (MI355X)
COUNTERS = os.path.join(ROOT, "profiles", "move_counters.json")
STORES = ("i16_tile", "bf16_tile", "f16_tile", "bf16", "f16", "f32")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes_per_step(store_bytes: int, static_bytes: int, midpoint: bool, random: bool) -> int:
    """Unique forcing values the reference stencil reads in one live substep
    (Particle.cpp:271-290, 352-500, 585-641, 709-739), both records counted:
      dynamic: zeta(X) 8, u 8, v 8, w 16, zeta(X') 8 [+ midpoint: zeta(X*) 8,
               u/v/w(X*) 32] [+ diffusion: AKt 16]
      static:  h(X) 4, dx+dy 2, depthIntervals 1, h(X') 4 [+ midpoint: h(X*) 4,
               dx+dy(X*) 2, depthIntervals 1]."""
    n_dyn = 48 + (40 if midpoint else 0) + (16 if random else 0)
    n_stat = 11 + (7 if midpoint else 0)
    return n_dyn * store_bytes + n_stat * static_bytes


STATE_BYTES = 96   # FP64 SoA: read id,k,j,i,health,age,time (56 B) + write k,j,i,health,age (40 B); the drift
                   # tag is touched only with the jibe on


def tune_sort(ctx, W, cfg, p0, path, at=12):
    """The cell-tile sort (wc_set_sort) the autotuner picks on a representative
    state: the population moved to `at` h (plumes dispersed), timed on the
    next interval; the population is then reloaded by the caller."""
    ctx.set_sort(0)
    ctx.set_particles(p0)
    for s in range(at):
        ctx.move_interval(cfg, float(W.times[s]), float(W.times[s + 1]), s, s + 1, path=path)
        ctx.remove_dead()
    tile, _ = ctx.autotune_sort(cfg, float(W.times[at]), float(W.times[at + 1]), at, at + 1, path=path)
    return tile


def run_timed(ctx, W, cfg, p0, steps, warmup, path, progress="", tile=0):
    """Warmup intervals, then `steps` timed intervals from t0 (records
    resident), with the cell-tile sort `tile` (0: none)."""
    ctx.set_sort(tile, 1)
    for w in range(warmup):
        ctx.set_particles(p0)
        ctx.move_interval(cfg, float(W.times[w]), float(W.times[w + 1]), w, w + 1, path=path)
        ctx.remove_dead()
        ctx.grid(counts=False, conc=False, mask_output=True)
        ctx.sync()
    ctx.set_particles(p0)
    ctx.sync()
    ctx.reset_kernel_time()
    ctx.reset_step_count()
    ctx.set_timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    for s in range(steps):
        ctx.move_interval(cfg, float(W.times[s]), float(W.times[s + 1]), s, s + 1, path=path)
        ctx.remove_dead()                 # synchronous (returns the survivor count)
        ctx.grid(counts=False, conc=False, mask_output=True)
        if s % 6 == 5:
            log(f"  {progress} interval {s + 1}/{steps}")
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.set_timing(False)
    move_ms, move_n = ctx.kernel_time(0)
    grid_ms, grid_n = ctx.kernel_time(1)
    sort_ms, sort_n = ctx.kernel_time(3)
    pack_ms, pack_n = ctx.kernel_time(4)
    ctx.set_sort(0)
    return dict(wall_s=wall, steps=ctx.step_count(), move_ms=move_ms, move_launches=move_n,
                grid_ms=grid_ms, grid_launches=grid_n, sort_tile=tile, sort_ms=sort_ms, sort_launches=sort_n,
                pack_ms=pack_ms, pack_launches=pack_n, alive=ctx.count())


def load_window(ctx, W, store, n_slots=None):
    ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask, n_slots=n_slots or W.T,
                 store=store)
    for t in range(n_slots or W.T):
        ctx.upload_record(t, *W.record(t))


def roofline_block(key, r, n_particles, store_bytes, static_bytes, midpoint, counters):
    """VALU-issue roofline of one move kernel (the binding resource), with the
    measured L2-to-memory traffic and the L2/MALL-effective bandwidth beside
    it.  `key` names the run (variant, integrator, cell-tile sort); the
    counters hold one value per interval from t0, and the run's timed
    intervals 0 .. steps-1 are averaged."""
    avg_s = r["move_ms"] / 1e3 / r["move_launches"]
    steps_per_launch = r["steps"] / r["move_launches"]
    eff = (steps_per_launch * algorithmic_bytes_per_step(store_bytes, static_bytes, midpoint, True) +
           n_particles * STATE_BYTES) / avg_s / 1e9
    c = (counters or {}).get(key)
    block = {"bound": "valu", "unit": "wave-instr/s", "peak": VALU_ISSUE_PEAK, "achieved": None, "frac": None,
             "avg_launch_ms": avg_s * 1e3, "effective_gbs": eff,
             "effective_gbs_note": "reference-stencil algorithmic bytes / launch time; served by L2 and the "
                                   "Infinity Cache, not an HBM fraction",
             "counters": None, "traffic": None, "hbm": None}
    n = r["move_launches"]
    if c and len(c["valu_per_interval"]) >= n:
        valu = float(np.mean(c["valu_per_interval"][:n]))
        achieved = valu / avg_s
        block.update(achieved=achieved, frac=achieved / VALU_ISSUE_PEAK,
                     counters={"key": key, "kernel": c["kernel"], "valu_per_dispatch": valu,
                               "intervals": n, "source": os.path.relpath(COUNTERS, ROOT)})
        if c.get("mem_bytes_per_interval"):
            t = float(np.mean(c["mem_bytes_per_interval"][:n]))
            hbm = t / avg_s / 1e9
            block.update(traffic=t, hbm={"achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                          "frac": hbm / HBM_PEAK_GBS,
                                          "what": "FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by its "
                                                  "reading for a plain copy of known size (L2-to-memory "
                                                  "traffic: HBM or the Infinity Cache) / launch time"})
    return block


def roms_records(W, t, mask_u, mask_v):
    """Record t of the window restated on ROMS's staggered grids: u at the
    u faces (mean of the two rho points beside each face), v at the v faces,
    fill values on land faces."""
    zeta, u, v, w, akt = W.record(t)
    uf = (0.5 * (u[:, :, :-1] + u[:, :, 1:])).astype(np.float32)
    vf = (0.5 * (v[:, :-1, :] + v[:, 1:, :])).astype(np.float32)
    uf[:, mask_u == 0] = np.float32(9.9692099683868690e+36)
    vf[:, mask_v == 0] = np.float32(9.9692099683868690e+36)
    return zeta, uf, vf, w, akt


def dropin_bench(wc, W, cfg, p0, steps):
    """The literal cudaMoveParticle replacement, wc_hip_move_particles
    (kernel.cu:370-382): particle_data AoS (64 B) and the whole window in the
    reference's device layout (zeta [T][eta][xi], u/v [T][s_rho][eta][xi],
    w/akt [T][s_w][eta][xi], FP32), one call per interval as Wacomm::run makes
    it (each call packs the interval's record pair, then runs the FP64 parity
    kernel over the pack).  Particle-steps are
    recovered from the ages (every particle is released at t0 and a live
    substep ages it by dti; the substep that kills a particle does not, so
    the count is a few in 1e5 below the move's own count)."""
    import ctypes as C
    L = wc.load_library()
    hip = C.CDLL("libamdhip64.so.7")               # the runtime the library already loaded
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    bufs = []

    def dev(a):
        a = np.ascontiguousarray(a)
        ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(ptr), max(1, a.nbytes)) == 0
        assert hip.hipMemcpy(ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0
        bufs.append(ptr)
        return ptr

    try:
        zeta, u, v, w, akt = W.stacked()
        f64 = lambda a: np.asarray(a, np.float64)
        ptrs = [dev(f64(W.times)), dev(f64(W.mask)), dev(f64(W.lonRad)), dev(f64(W.latRad)),
                dev(f64(W.depthIntervals)), dev(f64(W.h))] + [dev(a.astype(np.float32)) for a in (zeta, u, v, w, akt)]
        del zeta, u, v, w, akt
        p = np.ascontiguousarray(p0.astype(wc.PARTICLE_DTYPE))
        pd, cd = dev(p), dev(np.frombuffer(bytes(cfg), np.uint8))

        def call(t):
            rc = L.wc_hip_move_particles(cd, pd, t, W.T, W.s_w, W.s_rho, W.eta, W.xi, *ptrs, None, None, None, None,
                                         len(p), 1, 1, None)
            assert rc == wc.WC_OK, rc

        call(0)                                        # warm-up: first call builds the metric table
        assert hip.hipMemcpy(pd, p.ctypes.data_as(C.c_void_p), p.nbytes, 1) == 0
        assert hip.hipDeviceSynchronize() == 0
        t0 = time.perf_counter()
        for t in range(steps):
            call(t)
        assert hip.hipDeviceSynchronize() == 0
        wall = time.perf_counter() - t0
        out = np.empty_like(p)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), pd, out.nbytes, 2) == 0
        n_steps = int(np.rint(np.sum(out["age"] - p["age"]) / cfg.dti))
        return dict(value=n_steps / wall, unit="particle-steps/s", ms_per_interval=1e3 * wall / steps,
                    particle_steps=n_steps, intervals=steps, particles=len(p),
                    what="wc_hip_move_particles (cudaMoveParticle's argument list): AoS particle_data and the "
                         "window in the reference's device layout, one call per interval (pack of the record "
                         "pair, then the FP64 parity kernel)")
    finally:
        for b in bufs:
            hip.hipFree(b)


def roms_normalize_bench(wc, ctx, W, n_records):
    """ROMS normalisation of full-size records (1024x768, 11 levels) into
    forcing slots: device kernel time per record and the algorithmic bytes
    it moves, plus the host-to-device upload of the raw record."""
    m = W.mask
    mask_u = m[:, :-1] * m[:, 1:]
    mask_v = m[:-1, :] * m[1:, :]
    lon, lat = W.lonRad / 0.0174533, W.latRad / 0.0174533
    angle = np.full(m.shape, 0.3)
    ctx.set_roms_grid(np.linspace(-1.0, 0.0, W.s_w), lon, lat, W.h, m, mask_u, mask_v, angle=angle, n_slots=2)
    recs = [roms_records(W, t, mask_u, mask_v) for t in range(min(n_records, W.T))]
    ctx.upload_roms_record(0, *recs[0])            # warm-up
    ctx.sync()
    ctx.reset_kernel_time()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for q, r in enumerate(recs):
        ctx.upload_roms_record(q % 2, *r)
    wall = time.perf_counter() - t0
    ctx.set_timing(False)
    ms, n = ctx.kernel_time(2)
    plane, s_rho = W.eta * W.xi, W.s_rho
    pu, pv = mask_u.size, mask_v.size
    land = int((m == 0).sum())
    # reads: raw u, v, zeta, the three masks, {cos, sin} at wet points;
    # writes: U, V, and W, AKT, zeta at land points
    nbytes = (4 * s_rho * (pu + pv) + 4 * plane + plane + pu + pv + 16 * (plane - land) +
              8 * s_rho * plane + 4 * (2 * W.s_w + 1) * land)
    kernel_ms = ms / max(1, n)
    achieved = nbytes / (kernel_ms / 1e3) / 1e9
    return dict(records=n, kernel_ms=kernel_ms, algorithmic_bytes=nbytes, achieved_gbs=achieved,
                frac=achieved / HBM_PEAK_GBS, upload_ms_pcie_inclusive=1e3 * wall / len(recs))


def ensemble_bench(wc, ctx, W, cfg, steps):
    """The batched-ensemble workload: 256 release scenarios on the same grid
    and forcing (lognormal counts, median 40k, clipped to 10k-500k, ~16M in
    total; own source, Kh log-uniform in [0.1, 10] m2/s, release offset
    0-6 h), FP32 state, BF16 forcing, per-scenario surface counts every
    interval.  The launch is autotuned on the interval from 12 h, on the
    population moved there (every scenario released, plumes dispersed); the
    `steps` intervals then run once with direct launches and once as
    replays of the captured step, each from the same initial population."""
    from wacommplusplus_amd.synthetic import ensemble
    log("[ensemble] generating 256 scenarios")
    p, scen, sigma, counts = ensemble(W)
    load_window(ctx, W, "bf16")
    times = [float(W.times[q]) for q in range(steps + 1)]
    slots = list(range(steps + 1))
    ctx.ens_set(p, scen, sigma)
    # tune on a representative state: the population after 12 h (every
    # scenario released, plumes dispersed), on the next interval; the timed
    # runs below start again from the initial population
    tq = min(12, steps - 1)
    if tq > 0:
        ctx.ens_run(cfg, times[:tq + 1], slots[:tq + 1], use_graph=False, surface=False)
    best_ms = ctx.ens_autotune(cfg, times[tq], times[tq + 1], tq, tq + 1)
    shape = ctx.ens_get_launch()
    tile = ctx.ens_get_sort()[0]
    lds_chunk = ctx.ens_get_lds()[0]
    log(f"[ensemble] {len(p)} particles; autotuned block={shape[0]} ppt={shape[1]} occ={shape[2]} lds chunk={lds_chunk} "
        f"sort tile={tile} ({best_ms:.2f} ms on the interval from {tq} h)")
    out = {}
    for use_graph in (False, True):
        ctx.ens_set(p, scen, sigma)
        st = ctx.ens_run(cfg, times, slots, use_graph=use_graph, surface=True)
        key = "graph" if use_graph else "direct"
        out[key] = dict(value=st.particle_steps / st.wall_s, device_ms=st.device_ms, wall_s=st.wall_s,
                        particle_steps=int(st.particle_steps))
        log(f"[ensemble] {key}: {out[key]}")
    surf = ctx.ens_surface()
    return dict(metric="particle-steps/s", value=out["graph"]["value"], scenarios=len(sigma), particles=len(p),
                intervals=steps, launch=dict(block=shape[0], ppt=shape[1], occupancy_hint=shape[2], sort_tile=tile,
                                                 lds_chunk=lds_chunk),
                autotune_interval=[tq, tq + 1], autotune_ms=best_ms, graph=out["graph"], direct=out["direct"],
                surface_counts_last_interval=int(surf.sum()), dtype="f32 state, bf16 forcing",
                counts=dict(min=int(counts.min()), median=float(np.median(counts)), max=int(counts.max())))


def host_info():
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info["lscpu " + k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(ctx, W, cfg, p0, runs=3):
    """The oracle (FP64 C restatement of Particle::move, OpenMP
    schedule(dynamic,64) as Wacomm.cpp:501) on a bounded sample spanning the
    24 h: particles sampled from the release are moved on the GPU (FP64
    parity path, the same bits as the oracle) to 0, 6, 12 and 18 h, and one
    timed run moves every snapshot through its next interval on the host.
    All the host cores this job may use (OMP_PLACES=cores,
    OMP_PROC_BIND=close) and one core; midpoint and Euler; one warm-up and
    `runs` measured runs each, median and range reported
    (docs/performance-evaluation.md:7 of the reference)."""
    os.environ.setdefault("OMP_PLACES", "cores")
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_DYNAMIC", "false")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from runner import oracle_pair, cfg_to_oracle
    import pyoracle as po
    import wacommplusplus_amd as wc
    info = host_info()
    env_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    all_cores = min(info["affinity"], env_threads) if env_threads > 0 else info["affinity"]
    hours = (0, 6, 12, 18)
    per_core = 6000                     # particles per snapshot per thread: ~1 s per midpoint run
    n_max = per_core * all_cores
    stride = max(1, len(p0) // n_max)
    sample = p0[::stride][:n_max].copy()
    snaps = {}
    load_window(ctx, W, "f32", n_slots=W.T)
    ctx.set_particles(sample)
    for t in range(max(hours) + 1):
        if t in hours:
            snaps[t] = ctx.get_particles()
        ctx.move_interval(cfg, float(W.times[t]), float(W.times[t + 1]), t, t + 1, path="parity")
    windows = {t: oracle_pair(W, t, t + 1)[0] for t in hours}
    out = {"host": info, "threads_all": all_cores, "sample": None, "runs": runs}
    for integ in ("midpoint", "euler"):
        c = cfg.copy(integrator=wc.MIDPOINT if integ == "midpoint" else wc.EULER)
        oc = cfg_to_oracle(c)
        for label, threads in (("all_cores", all_cores), ("one_core", 1)):
            n = per_core * threads
            rates = []
            steps = 0
            for r in range(runs + 1):
                steps = 0
                t0 = time.perf_counter()
                for t in hours:
                    q = snaps[t][::max(1, all_cores // threads)][:n].copy()     # spread over the sources
                    alive = int((q["health"] > 0).sum())
                    po.move(windows[t], oc, 0, q, nthreads=threads)
                    steps += alive * int(round(3600.0 / c.dti))
                wall = time.perf_counter() - t0
                if r > 0:
                    rates.append(steps / wall)
            rates.sort()
            out[f"{integ}_{label}"] = {"median": rates[len(rates) // 2], "min": rates[0], "max": rates[-1],
                                       "threads": threads, "particles_per_snapshot": n,
                                       "particle_steps_per_run": steps}
            log(f"[cpu] {integ} {label} ({threads} threads): median {rates[len(rates) // 2]:.4g} particle-steps/s")
    head = out["midpoint_all_cores"]
    return dict(value=head["median"], unit="particle-steps/s", cores=head["threads"], kind="port",
                sample=f"every {stride}th particle of the 4M release ({per_core} per thread), moved on the GPU to "
                       f"0, 6, 12 and 18 h; one run moves each snapshot through its next interval with the oracle "
                       f"(FP64, OpenMP dynamic,64, midpoint, diffusion on); {runs} runs after a warm-up, median",
                detail=out)


def precision_vs_f64(ctx, W, p, steps, cfg):
    """The mixed-precision path against the FP64 path after `steps` hours,
    diffusion off, each particle offset by N(0, 0.25^2) cells: per-particle
    max(|dk|, |dj|, |di|) over the particles alive on both paths, over the
    closure-free ones (no shore or open-boundary closure and no midpoint
    fallback at land on either path; the rest are held on coast node lines by
    the reflection closure), and over the open-water ones (closure-free, and
    no stencil with a land node on either path, where the bilinear operator
    jumps to the land zeros).  With diffusion on, the relative L2 of the
    count grid against the FP64 path and against a second FP64 seed."""
    import wacommplusplus_amd as wc
    pin = wc.EVENT["shore"] | wc.EVENT["domain"] | wc.EVENT["fallback"]
    coastal = pin | wc.EVENT["coast"]
    det = cfg.copy(random=False)

    def run(path, store, c, parts, events):
        load_window(ctx, W, store)
        ctx.set_particles(parts)
        ctx.set_event_tracking(events)
        for s in range(steps):
            ctx.move_interval(c, float(W.times[s]), float(W.times[s + 1]), s, s + 1, path=path)
        out = ctx.get_particles(), (ctx.get_events() if events else None)
        ctx.set_event_tracking(False)
        return out

    ref, ev_ref = run("parity", "f32", det, p, True)
    rnd = p[::10]
    run("parity", "f32", cfg, rnd, False)
    g0 = ctx.grid().astype(np.float64)
    run("parity", "f32", cfg.copy(randomSeed=1234), rnd, False)
    g1 = ctx.grid().astype(np.float64)
    nrm = np.linalg.norm(g0)
    res = {"hours": steps, "diffusion": "off", "particles": len(p),
           "release": "50 sources, per-particle N(0, 0.25^2)-cell offsets",
           "closure_free": "no WC_EVENT_SHORE, DOMAIN or FALLBACK on either path",
           "open_water": "closure-free and no WC_EVENT_COAST (a stencil with a land node) on either path",
           "conc_l2": {"diffusion": "on", "particles": len(rnd),
                       "f64_seed_to_seed": float(np.linalg.norm(g1 - g0) / nrm)}}
    for store in STORES:
        a, ev = run("fast", store, det, p, True)
        both = (a["health"] > 0) & (ref["health"] > 0)
        d = np.max(np.stack([np.abs(a[f] - ref[f]) for f in ("k", "j", "i")]), axis=0)
        free = both & (((ev | ev_ref) & pin) == 0)
        open_water = both & (((ev | ev_ref) & coastal) == 0)
        r = {}
        for name, sel in (("alive", both), ("closure_free", free), ("open_water", open_water)):
            x = d[sel]
            q = np.quantile(x, [0.5, 0.9, 0.99])
            r[name] = dict(fraction=float(sel.mean()), p50_cells=float(q[0]), p90_cells=float(q[1]),
                           p99_cells=float(q[2]), max_cells=float(x.max()), rms_cells=float(np.sqrt(np.mean(x * x))))
        run("fast", store, cfg, rnd, False)
        r["conc_l2"] = float(np.linalg.norm(ctx.grid().astype(np.float64) - g0) / nrm)
        res[store] = r
        log(f"[precision] {store}: {r}")
    return res


EULER_VARIANTS = ("f64", "i16_tile", "bf16")
LDS_TILING = (30, 2)      # chunk of substeps per staged box, halo cells


def profile_runs(ctx, W, cfg, p0, variants, tiles, args):
    """The runs the committed move-kernel counters come from: each variant
    unsorted and at every listed cell-tile sort, midpoint and (for
    EULER_VARIANTS) Euler, `warmup` + `steps` launches each, in the order
    printed.  The counters are per launch, so their per-interval values
    serve any run of the same variant, integrator and sort."""
    import wacommplusplus_amd as wc
    runs = []
    for path, store in variants:
        key = "f64" if path == "parity" else store
        load_window(ctx, W, store)
        for euler in ((False, True) if not args.no_euler and key in EULER_VARIANTS else (False,)):
            c = cfg.copy(integrator=wc.EULER) if euler else cfg
            for tile in [0] + tiles:
                name = f"{key}{'_euler' if euler else ''}@t{tile}"
                r = run_timed(ctx, W, c, p0, args.steps, args.warmup, path, progress=name, tile=tile)
                runs.append(dict(key=name, variant=key, path=path, store=store, sort_tile=tile,
                                 warmup_launches=args.warmup, timed_launches=r["move_launches"],
                                 move_kernel_ms=r["move_ms"] / r["move_launches"]))
                log(f"[{name}] {r}")
    print(json.dumps({"profile_runs": runs, "steps": args.steps, "warmup": args.warmup,
                      "particles": args.particles}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=4_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-euler", action="store_true", help="skip the Euler-integrator timings")
    ap.add_argument("--no-sort", action="store_true", help="no cell-tile sort (skip its autotuning)")
    ap.add_argument("--no-lds", action="store_true", help="skip the LDS-staged BF16 move")
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in entry point's rate")
    ap.add_argument("--no-pcie", action="store_true", help="skip the upload-inclusive rate")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble workload")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-reconstruction workload")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the precision comparison (its diffusion-off runs launch the move kernels "
                         "on other populations, which a kernel-trace average would mix in)")
    ap.add_argument("--sort-tiles", default=None,
                    help="counter-collection runs: comma-separated cell-tile sorts; every variant is timed "
                         "unsorted and at each of them (midpoint, and Euler where reported), without the "
                         "autotuner, and one JSON line lists the runs in launch order")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU only: the particle ranges of several "
                                   "GPUs move independently and only the count frames are summed, which "
                                   "is not run here", "n_gpus": args.gpus}), flush=True)
        sys.exit(2)
    assert 1 <= args.steps <= 24, "the window holds 24 intervals"
    import wacommplusplus_amd as wc
    from wacommplusplus_amd.synthetic import full_window, release

    log("generating the 1024x768x11 window (25 records) and the 4M release")
    W = full_window()
    p0 = release(W, args.particles, 50)
    cfg = wc.default_config(random=True, sigma=wc.sigma_from_kh(1.0, 30.0), crid=1.0)
    counters = None
    if os.path.exists(COUNTERS):
        with open(COUNTERS) as f:
            counters = json.load(f).get("kernels")

    results = {}
    variants = [("parity", "f32")] + ([] if args.no_variants else [("fast", s) for s in STORES])
    ctx = wc.Context(0)
    if args.sort_tiles is not None:
        profile_runs(ctx, W, cfg, p0, variants, [int(t) for t in args.sort_tiles.split(",") if t], args)
        ctx.close()
        return
    for path, store in variants:
        key = "f64" if path == "parity" else store
        log(f"[{key}] uploading {W.T} records ({store} storage)")
        load_window(ctx, W, store)
        log(f"[{key}] timing {args.steps} intervals ({args.warmup} warmup)")
        results[key + "_unsorted"] = run_timed(ctx, W, cfg, p0, args.steps, args.warmup, path, progress=key)
        log(f"[{key} unsorted] {results[key + '_unsorted']}")
        tile = 0 if args.no_sort else tune_sort(ctx, W, cfg, p0, path)
        log(f"[{key}] autotuned cell-tile sort: {tile}")
        results[key] = (run_timed(ctx, W, cfg, p0, args.steps, args.warmup, path, progress=key, tile=tile)
                        if tile else results[key + "_unsorted"])
        log(f"[{key}] {results[key]}")
        if not args.no_euler and key in EULER_VARIANTS:
            e = run_timed(ctx, W, cfg.copy(integrator=wc.EULER), p0, args.steps, args.warmup, path,
                          progress=key + " euler", tile=tile)
            results[key + "_euler"] = e
            log(f"[{key} euler] {e}")
        if key == "bf16" and not args.no_lds:
            # the same move with the forcing staged in LDS (wc_set_lds), at the autotuned cell-tile order
            ctx.set_lds(*LDS_TILING)
            e = run_timed(ctx, W, cfg, p0, args.steps, args.warmup, path, progress=key + " lds", tile=tile)
            ctx.set_lds(0)
            results[key + "_lds"] = e
            log(f"[{key} lds {LDS_TILING}] {e}")

    precision = None
    if not args.no_variants and not args.no_precision:
        precision = precision_vs_f64(ctx, W, release(W, args.particles, 50, spread=0.25), args.steps, cfg)

    # upload-inclusive rate: a three-slot ring, one record uploaded per
    # interval from host memory -- the record after next, copied on the copy
    # stream while the move of the current interval runs
    pcie = None
    if not args.no_pcie:
        load_window(ctx, W, "f32", n_slots=3)
        ctx.set_particles(p0)
        ctx.reset_step_count()
        ctx.sync()
        t0 = time.perf_counter()
        ctx.upload_record(0, *W.record(0))
        ctx.upload_record(1, *W.record(1))
        for s in range(args.steps):
            ctx.move_interval(cfg, float(W.times[s]), float(W.times[s + 1]), s % 3, (s + 1) % 3, path="parity")
            if s + 2 < W.T:
                ctx.upload_record((s + 2) % 3, *W.record(s + 2))
            ctx.remove_dead()
            ctx.grid(counts=False, mask_output=True)
        ctx.sync()
        pcie = ctx.step_count() / (time.perf_counter() - t0)
        log(f"[pcie] {pcie:.4g} particle-steps/s")
    roms = roms_normalize_bench(wc, ctx, W, args.steps)
    log(f"[roms] {roms}")
    dropin = None
    if not args.no_dropin:
        ctx.set_particles(wc.make_particles(0))            # release the context's copy first
        dropin = dropin_bench(wc, W, cfg, p0, args.steps)
        log(f"[drop-in] {dropin}")
    ens = None if args.no_ensemble else ensemble_bench(wc, ctx, W, cfg, args.steps)
    learned = None
    if not args.no_learned:
        from wacommplusplus_amd import reconstruct
        log("[learned] 512 training + 128 held-out pairs (1M-particle targets, 50k inputs, 24 h)")
        load_window(ctx, W, "bf16")
        learned = reconstruct.run(ctx, W, n_train=512, n_test=128, epochs=40, log=log)
        log(f"[learned] {learned['held_out']} {learned['equal_gpu_ms']}")
    cpu = None if args.no_cpu else cpu_baseline(ctx, W, cfg, p0)
    ctx.close()

    main_r = results["f64"]
    value = main_r["steps"] / main_r["wall_s"]
    roof = roofline_block(f"f64@t{main_r['sort_tile']}", main_r, args.particles, 4, 8, True, counters)
    roof["kernel"] = "move_parity_kernel<PackedSampler, SoAAccess>"
    roof["sort_tile"] = main_r["sort_tile"]
    roof["algorithmic_bytes_per_particle_step"] = algorithmic_bytes_per_step(4, 8, True, True)
    variants_out = {}
    for key, r in results.items():
        if key == "f64":
            continue
        store = key.replace("_euler", "").replace("_unsorted", "").replace("_lds", "")
        euler = key.endswith("_euler")
        ckey = f"{key.replace('_unsorted', '')}@t{r['sort_tile']}"
        if store == "f64":
            b = roofline_block(ckey, r, args.particles, 4, 8, not euler, counters)
        else:
            b = roofline_block(ckey, r, args.particles, 4 if store == "f32" else 2, 4, not euler, counters)
        variants_out[key] = dict(value=r["steps"] / r["wall_s"], ms_per_step=1e3 * r["wall_s"] / args.steps,
                                 move_kernel_ms=r["move_ms"] / r["move_launches"],
                                 grid_kernel_ms=r["grid_ms"] / max(1, r["grid_launches"]),
                                 pack_ms_per_interval=r["pack_ms"] / max(1, r["pack_launches"]),
                                 sort_tile=r["sort_tile"], sort_ms_per_interval=r["sort_ms"] / args.steps,
                                 integrator="euler" if euler else "midpoint",
                                 path="parity (FP64)" if store == "f64" else f"fast ({store} storage)",
                                 roofline=b)
        if key.endswith("_lds"):
            variants_out[key]["lds"] = dict(chunk=LDS_TILING[0], halo=LDS_TILING[1],
                                            what="forcing boxes staged in LDS per chunk of substeps (wc_set_lds); "
                                                 "results identical to the plain move")
    line = {
        "metric": "particle-steps/s at 4M (BF16 vs FP64 forcing); conc. L2 err at equal GPU-ms",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * main_r["wall_s"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"full-size plain method: 1024x768x11 sigma grid at 150 m, 64-mode forcing, ~30% land, "
                        f"4M particles from 50 sources, Kh=1 m2/s, dt=30 s, midpoint, {args.steps} h "
                        f"({args.steps} intervals); FP64 state, FP32-stored forcing (the reference's storage), "
                        f"records resident in HBM",
            "particles": args.particles, "grid": [11, 1024, 768], "substeps_per_step": 120,
            "path": "parity (FP64)", "records": "resident in HBM (upload-inclusive rate in pcie_inclusive_value)",
            "cell_tile_sort": main_r["sort_tile"],
            "sort_autotuned": not args.no_sort,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "variants": variants_out,
        "north_star_kernel": "fast path, 16-bit storage: i16_tile (block-scaled int16) for the drift bar, bf16 "
                             "for throughput (variants.i16_tile, variants.bf16)",
        "precision_vs_f64": precision,
        "pcie_inclusive_value": pcie,
        "dropin": dropin,
        "roms_normalize": roms,
        "ensemble": ens,
        "learned_reconstruction": learned,
        "grid_kernel_ms": main_r["grid_ms"] / max(1, main_r["grid_launches"]),
        "pack_ms_per_interval": main_r["pack_ms"] / max(1, main_r["pack_launches"]),
        "sort_ms_per_interval": main_r["sort_ms"] / args.steps,
        "alive_after": main_r["alive"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark of the particle-move + gridding path on one MI355X.

Workload (the full-size plain-method configuration): 1024(eta) x 768(xi) x 11
sigma-level grid at 150 m with ~30 % land, u/v from 64 random Fourier modes
(5-40 km, rms 0.3 m/s), w from continuity, 25 hourly records; 4,000,000
particles released at t0 from 50 random ocean point sources at 1 m depth;
Kh = 1 m2/s (sigma = sqrt(2 Kh dti)), AKt = 1e-4 m2/s, dti = 30 s, explicit
midpoint, default closures.  Synthetic data generated in memory.

A step is one forcing interval of the hot path for the whole population:
move (120 substeps) -> remove dead -> grid, i.e. one pass of the interval
loop of Wacomm::run.  All forcing records are resident in HBM before the
timed region.  The headline value is the FP64-state parity path (FP32-stored
forcing, as the reference stores it); the mixed-precision path with BF16,
FP16 and FP32 storage is reported alongside.

    python bench.py [--gpus 1] [--steps 24] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4     # wave64 VALU instructions per second (MI355X)


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


def run_timed(wc, ctx, W, cfg, p0, steps, warmup, path, progress=""):
    """Warmup intervals, then `steps` timed intervals from t0 (records resident)."""
    for w in range(warmup):
        ctx.set_particles(p0)
        ctx.move_interval(cfg, float(W.times[w]), float(W.times[w + 1]), w, w + 1, path=path)
        ctx.remove_dead()
        ctx.grid(counts=False, conc=False)
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
        ctx.grid(counts=False, conc=False)
        if s % 6 == 5:
            log(f"  {progress} interval {s + 1}/{steps}")
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.set_timing(False)
    move_ms, move_n = ctx.kernel_time(0)
    grid_ms, grid_n = ctx.kernel_time(1)
    return dict(wall_s=wall, steps=ctx.step_count(), move_ms=move_ms, move_launches=move_n,
                grid_ms=grid_ms, grid_launches=grid_n, alive=ctx.count())


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
    ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask, n_slots=W.T, store="bf16")
    for t in range(W.T):
        ctx.upload_record(t, *W.record(t))
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


def cpu_baseline(W, cfg, p0, n_sample=100_000, n_intervals=4):
    """The oracle (FP64 C restatement of Particle::move, OpenMP dynamic,64 as
    Wacomm.cpp:501) on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from runner import oracle_pair, cfg_to_oracle
    import pyoracle as po
    stride = max(1, len(p0) // n_sample)
    p = p0[::stride].copy()
    ocfg = cfg_to_oracle(cfg)
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    windows = [oracle_pair(W, t, t + 1)[0] for t in range(n_intervals)]
    steps = 0
    t0 = time.perf_counter()
    for t in range(n_intervals):
        alive_before = p["health"] > 0
        po.move(windows[t], ocfg, 0, p, nthreads=cores)
        steps += int(alive_before.sum())
    wall = time.perf_counter() - t0
    # every alive particle takes the 120 substeps of the interval (single pulse at t0)
    psteps = steps * int(round(3600.0 / cfg.dti))
    return dict(value=psteps / wall, unit="particle-steps/s", cores=cores, kind="port",
                sample=f"{len(p)} of the 4M particles (every {stride}th), first {n_intervals} intervals "
                       f"({psteps:.3g} particle-steps, {wall:.1f} s wall on {cores} threads)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=4_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble workload")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-reconstruction workload")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the precision comparison (its diffusion-off runs launch the move kernels "
                         "on other populations, which a kernel-trace average would mix in)")
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
    n_slots = W.T

    results = {}
    variants = [("parity", "f32")] + ([] if args.no_variants else [("fast", "bf16"), ("fast", "f16"),
                                                                   ("fast", "f32")])
    ctx = wc.Context(0)
    for path, store in variants:
        key = "f64" if path == "parity" else store
        log(f"[{key}] uploading {n_slots} records ({store} storage)")
        ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask, n_slots=n_slots,
                     store=store)
        for t in range(n_slots):
            ctx.upload_record(t, *W.record(t))
        log(f"[{key}] timing {args.steps} intervals ({args.warmup} warmup)")
        results[key] = run_timed(wc, ctx, W, cfg, p0, args.steps, args.warmup, path, progress=key)
        log(f"[{key}] {results[key]}")

    # precision of the mixed-precision path against the FP64 path:
    #  (a) per particle, diffusion off, whole population, `steps` intervals;
    #  (b) the count grid after `steps` intervals, diffusion on, against the
    #      difference between two FP64 runs with different seeds
    precision = {"per_particle": {}, "conc_l2": {}}
    if not args.no_variants and not args.no_precision:
        det = cfg.copy(random=False)
        pp = release(W, args.particles, 50, spread=0.25)      # distinct trajectories
        finals, grids = {}, {}
        for path, store in [("parity", "f32"), ("fast", "bf16"), ("fast", "f16"), ("fast", "f32")]:
            key = "f64" if path == "parity" else store
            ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask, n_slots=n_slots,
                         store=store)
            for t in range(n_slots):
                ctx.upload_record(t, *W.record(t))
            runs = [("det", det, pp), ("rnd", cfg, pp[::10])]
            if key == "f64":
                runs.append(("rnd2", cfg.copy(randomSeed=1234), pp[::10]))
            for tag, c, parts in runs:
                ctx.set_particles(parts)
                for s in range(args.steps):
                    ctx.move_interval(c, float(W.times[s]), float(W.times[s + 1]), s, s + 1, path=path)
                if tag == "det":
                    finals[key] = ctx.get_particles()
                else:
                    grids[(key, tag)] = ctx.grid().astype(np.float64)
            log(f"[precision] {key} done")
        ref = finals.pop("f64")
        for key, a in finals.items():
            both = (a["health"] > 0) & (ref["health"] > 0)
            d = np.max(np.stack([np.abs(a[f][both] - ref[f][both]) for f in ("k", "j", "i")]), axis=0)
            q = np.quantile(d, [0.5, 0.9, 0.99])
            precision["per_particle"][key] = dict(p50_cells=float(q[0]), p90_cells=float(q[1]),
                                                  p99_cells=float(q[2]), max_cells=float(d.max()),
                                                  both_alive=float(both.mean()))
        g0 = grids[("f64", "rnd")]
        nrm = np.linalg.norm(g0)
        precision["conc_l2"]["f64_seed_to_seed"] = float(np.linalg.norm(grids[("f64", "rnd2")] - g0) / nrm)
        for key in ("bf16", "f16", "f32"):
            precision["conc_l2"][key] = float(np.linalg.norm(grids[(key, "rnd")] - g0) / nrm)
        log(f"[precision] {precision}")

    # PCIe-inclusive rate: a three-slot ring, one record uploaded per
    # interval from host memory -- the record after next, copied on the
    # copy stream while the move of the current interval runs
    pcie = None
    if not args.no_variants:
        ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask, n_slots=3, store="f32")
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
            ctx.grid(counts=False)
        ctx.sync()
        pcie = ctx.step_count() / (time.perf_counter() - t0)
        log(f"[pcie] {pcie:.4g} particle-steps/s")
    roms = roms_normalize_bench(wc, ctx, W, args.steps)
    log(f"[roms] {roms}")
    ens = None if args.no_ensemble else ensemble_bench(wc, ctx, W, cfg, args.steps)
    learned = None
    if not args.no_learned:
        from wacommplusplus_amd import reconstruct
        log("[learned] 512 training + 128 held-out pairs (1M-particle targets, 50k inputs, 24 h)")
        ctx.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask, n_slots=W.T, store="bf16")
        for t in range(W.T):
            ctx.upload_record(t, *W.record(t))
        learned = reconstruct.run(ctx, W, n_train=512, n_test=128, epochs=40, log=log)
        log(f"[learned] {learned['held_out']} {learned['equal_gpu_ms']}")
    ctx.close()

    main_r = results["f64"]
    value = main_r["steps"] / main_r["wall_s"]
    bps = algorithmic_bytes_per_step(4, 8, True, True)
    avg_launch_s = main_r["move_ms"] / 1e3 / main_r["move_launches"]
    steps_per_launch = main_r["steps"] / main_r["move_launches"]
    bytes_per_launch = steps_per_launch * bps + args.particles * STATE_BYTES
    achieved = bytes_per_launch / avg_launch_s / 1e9
    traffic = None
    valu = None
    tpath = os.path.join(ROOT, "profiles", "move_parity_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            prof = json.load(f)
        traffic = prof.get("hbm_bytes_per_launch")
        valu = prof.get("valu_per_wave_substep")
    # instruction issue: wave-level VALU instructions per second against the
    # chip's issue rate (one wave64 VALU instruction per SIMD every 4 cycles:
    # 256 CUs x 4 SIMDs x 2.4 GHz / 4, the rate behind the 78.6 TFLOP/s FP64
    # vector peak = that x 64 lanes x 2 flop per FMA)
    issue = None
    if valu:
        wave_substeps = steps_per_launch / 64.0
        issue_rate = valu * wave_substeps / avg_launch_s
        issue = {"what": "VALU instructions per wave-substep (rocprofv3 SQ_INSTS_VALU) x wave-substeps per launch "
                         "/ the launch time measured here",
                 "valu_per_wave_substep": valu, "achieved": issue_rate, "peak": VALU_ISSUE_PEAK,
                 "unit": "wave-instr/s", "frac": issue_rate / VALU_ISSUE_PEAK}

    cpu = None if args.no_cpu else cpu_baseline(W, cfg, p0)

    variants_out = {}
    for key, r in results.items():
        if key == "f64":
            continue
        b = algorithmic_bytes_per_step(2 if key in ("bf16", "f16") else 4, 4, True, True)
        a_s = r["move_ms"] / 1e3 / r["move_launches"]
        spl = r["steps"] / r["move_launches"]
        variants_out[key] = dict(value=r["steps"] / r["wall_s"], ms_per_step=1e3 * r["wall_s"] / args.steps,
                                 move_kernel_ms=r["move_ms"] / r["move_launches"],
                                 achieved_gbs=(spl * b + args.particles * STATE_BYTES) / a_s / 1e9)
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
            "workload": "full-size plain method: 1024x768x11 sigma grid at 150 m, 64-mode forcing, ~30% land, "
                        "4M particles from 50 sources, Kh=1 m2/s, dt=30 s, midpoint, 24 h; FP64 state, "
                        "FP32-stored forcing (the reference's storage), records resident in HBM",
            "particles": args.particles, "grid": [11, 1024, 768], "substeps_per_step": 120,
            "path": "parity (FP64)",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "move_parity_kernel<SoAAccess>",
            "avg_launch_ms": avg_launch_s * 1e3, "algorithmic_bytes_per_particle_step": bps,
            "issue": issue,
        },
        "cpu_baseline": cpu,
        "variants": variants_out,
        "precision_vs_f64": {"hours": args.steps,
                             "per_particle": {"diffusion": "off", "particles": args.particles,
                                              "release": "50 sources, per-particle N(0, 0.25^2)-cell offsets",
                                              **precision["per_particle"]},
                             "conc_l2": {"diffusion": "on", "particles": args.particles // 10,
                                         "what": "relative L2 of the count grid vs the FP64 path",
                                         **precision["conc_l2"]}},
        "pcie_inclusive_value": pcie,
        "roms_normalize": roms,
        "ensemble": ens,
        "learned_reconstruction": learned,
        "grid_kernel_ms": main_r["grid_ms"] / max(1, main_r["grid_launches"]),
        "alive_after": main_r["alive"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Learned concentration reconstruction: training pairs generated by the
method itself, training and evaluation of the network (wc_unet_*).

Pairs.  For each pair a release point (a random ocean cell), a horizontal
diffusivity Kh (log-uniform) and a particle count N are drawn.  One batched
ensemble run moves the N particles; every `sub_every`-th particle (by id)
belongs to a second scenario, so the same run yields the surface-layer
histogram of the n = N / sub_every subsample (the input) and of all N
particles (the target).  Draws are keyed by particle id, so the subsample
moves exactly as it would alone.  Training images are crops around the
few-particle histogram's centre of mass; held-out images may be the whole
surface layer (crop=None).

Network input and target are variance-stabilised: x = sqrt(sub_every * few)
/ scale, t = sqrt(all) / scale (the square root of a Poisson count has
roughly constant variance); the reconstruction is (max(y, 0) * scale)^2.
Plain gridding -- the fallback -- is sub_every * few.

Loss.  The plain loss is the mean squared error in those units.  The error
that matters is the relative L2 norm in counts; to first order a unit of
error at a pixel of target t costs 4 s^4 t^2 there, over |c|^2 = s^4 sum t^4
for the image.  loss_weight="count" trains with per-pixel weights
t_b^2 / mean(t_b^4) (t_b: the target box-filtered over 3 x 3, so that its
noise does not steer the weights), plus a floor of 10 % of the image's mean,
scaled to a mean of one over the training set.

Extra input planes (`extras`) carry what is known at reconstruction time
besides the histogram: "kh", the run's horizontal diffusivity as a constant
plane log10(Kh); "mask", the crop's ocean mask (1 wet, 0 land).

Errors are relative L2 norms against the N-particle histogram.  That target
has its own sampling noise, estimated from the Poisson variance of a
histogram (E|h - lambda|^2 = sum lambda ~ sum h): e_target = sqrt(sum h) /
|h|.  Errors against the unknown noise-free field are then estimated as
sqrt(e^2 - e_target^2).
"""
from __future__ import annotations

import math
import time

import numpy as np

from ._lib import Context, EnsSource, UNet
from .synthetic import ForcingWindow, depth_to_k
from .types import default_config


def ocean_cells(W: ForcingWindow, n: int, margin: int, rng):
    wet = W.mask > 0
    ok = wet[:-1, :-1] & wet[1:, :-1] & wet[:-1, 1:] & wet[1:, 1:]
    ok[:margin, :] = False
    ok[-margin:, :] = False
    ok[:, :margin] = False
    ok[:, -margin:] = False
    cells = np.argwhere(ok)
    pick = cells[rng.choice(len(cells), n, replace=False)]
    off = rng.uniform(0.05, 0.95, (n, 2))
    return pick[:, 0] + off[:, 0], pick[:, 1] + off[:, 1]


def generate_pairs(ctx: Context, W: ForcingWindow, n_pairs: int, n_target: int = 1_000_000, sub_every: int = 20,
                   hours: int = 24, crop: int | None = 128, seed: int = 5489, kh_range=(0.1, 10.0),
                   dti: float = 30.0, depth_m: float = 1.0, batch_pairs: int = 128, frames: int = 1, log=None,
                   launch=(128, 2, 3, 4)):
    """Surface-layer histograms (few, all) of n_pairs releases after `hours`,
    cropped to crop x crop around the centre of mass of the few-particle
    histogram (what is known at reconstruction time; the release point when
    it is empty), or the whole eta x xi surface with crop=None.  launch:
    (block, particles per work-item, occupancy hint, sort tile) of the
    ensemble move, or None for the context's current one.  With frames > 1 the few-particle histograms of the
    frames - 1 hours before are kept too (a low-count run writes hourly
    frames), cropped at the same place.  The context must hold the window's
    grid and records (wc_set_grid with n_slots >= hours + 1).  Returns
    dict(few: uint32 [n_pairs][crop][crop], or [n_pairs][frames][crop][crop]
    with frame 0 at `hours`, frame f at hours - f; all: uint32
    [n_pairs][crop][crop]; mask: uint8 ocean mask of each crop; kh, j, i,
    origin [n_pairs][2], seconds,
    particle_steps)."""
    rng = np.random.default_rng(seed)
    full = crop is None
    ch, cw = (W.eta, W.xi) if full else (crop, crop)
    js, is_ = ocean_cells(W, n_pairs, (128 if full else crop) // 2 + 2, rng)
    kh = np.exp(rng.uniform(math.log(kh_range[0]), math.log(kh_range[1]), n_pairs))
    sigma = np.sqrt(2.0 * kh * dti)
    k0 = depth_to_k(depth_m, float(W.h[W.mask > 0].mean()), W.s_w)
    cfg = default_config(random=True, sigma=1.0, dti=dti)
    times = [float(W.times[q]) for q in range(hours + 1)]
    slots = list(range(hours + 1))
    few = np.zeros((n_pairs, frames, ch, cw), np.uint32)
    allc = np.zeros((n_pairs, ch, cw), np.uint32)
    wet = np.zeros((n_pairs, ch, cw), np.uint8)
    origin = np.zeros((n_pairs, 2), np.int64)
    jj, ii = np.arange(W.eta, dtype=np.float64), np.arange(W.xi, dtype=np.float64)
    t0 = time.perf_counter()
    steps = 0
    for b0 in range(0, n_pairs, batch_pairs):
        nb = min(batch_pairs, n_pairs - b0)
        sources = [EnsSource(k=k0, j=float(js[b0 + q]), i=float(is_[b0 + q]), spread=0.25, time=times[0],
                             count=n_target, scenario=2 * q + 1, sub_scenario=2 * q, sub_every=sub_every)
                   for q in range(nb)]
        sg = np.repeat(sigma[b0:b0 + nb], 2)
        ctx.ens_release(sources, sg, seed=seed + 7919 * (b0 + 1))
        if launch is not None:
            ctx.ens_set_launch(*launch[:3])
            ctx.ens_set_sort(launch[3], 1)
        # run to hours - (frames - 1), then one hour at a time, keeping the
        # surface counts after each of the last `frames` hours
        first = hours - (frames - 1)
        st = ctx.ens_run(cfg, times[:first + 1], slots[:first + 1], use_graph=True, surface=True)
        steps += st.particle_steps
        ms = st.device_ms
        kept = [ctx.ens_surface()]
        for hh in range(first, hours):
            st = ctx.ens_run(cfg, times[hh:hh + 2], slots[hh:hh + 2], use_graph=True, surface=True)
            steps += st.particle_steps
            ms += st.device_ms
            kept.append(ctx.ens_surface())
        kept = kept[::-1]                                 # frame 0 = the last hour
        surf = kept[0]
        for q in range(nb):
            fs = surf[2 * q].astype(np.float64)
            tot = fs.sum()
            cj = (fs.sum(1) @ jj) / tot if tot > 0 else js[b0 + q]
            ci = (fs.sum(0) @ ii) / tot if tot > 0 else is_[b0 + q]
            r0 = 0 if full else int(np.clip(round(cj) - ch // 2, 0, W.eta - ch))
            c0 = 0 if full else int(np.clip(round(ci) - cw // 2, 0, W.xi - cw))
            origin[b0 + q] = (r0, c0)
            wet[b0 + q] = W.mask[r0:r0 + ch, c0:c0 + cw] > 0
            f = surf[2 * q, r0:r0 + ch, c0:c0 + cw]
            for fr in range(frames):
                few[b0 + q, fr] = kept[fr][2 * q, r0:r0 + ch, c0:c0 + cw]
            allc[b0 + q] = f + surf[2 * q + 1, r0:r0 + ch, c0:c0 + cw]
        if log:
            log(f"  pairs {b0 + nb}/{n_pairs}: {steps:.3g} particle-steps so far, last batch {ms:.0f} ms")
    return dict(few=few[:, 0] if frames == 1 else few, all=allc, mask=wet, kh=kh, j=js, i=is_, origin=origin,
                seconds=time.perf_counter() - t0, particle_steps=steps, n_target=n_target, sub_every=sub_every,
                hours=hours, frames=frames, full=full)


EXTRAS = ("kh", "mask")


def extra_planes(data, extras):
    """The extra input planes of `data` ([n][len(extras)][h][w] float32), or
    None without extras."""
    if not extras:
        return None
    n, h, w = data["all"].shape
    planes = []
    for e in extras:
        if e == "kh":
            planes.append(np.broadcast_to(np.log10(data["kh"]).astype(np.float32)[:, None, None], (n, h, w)))
        elif e == "mask":
            planes.append(data["mask"].astype(np.float32))
        else:
            raise ValueError(f"unknown extra input {e!r} (known: {', '.join(EXTRAS)})")
    return np.ascontiguousarray(np.stack(planes, 1))


def to_net(few, allc, sub_every, scale, extra=None):
    """Network input (histogram frames, then the extra planes, as channels)
    and target, variance-stabilised."""
    x = (np.sqrt(sub_every * few.astype(np.float64)) / scale).astype(np.float32)
    if extra is not None:
        x = np.concatenate([x if x.ndim == 4 else x[:, None], extra], axis=1)
    t = None if allc is None else np.sqrt(allc.astype(np.float64)) / scale
    return x, (None if t is None else t.astype(np.float32))


def loss_weights(T, kind="count"):
    """Per-pixel loss weights for targets T [n][h][w] (variance-stabilised
    units): "count" (above), or "core", (t_b^2 + m2) / (2 m2) with m2 the
    image's mean t_b^2 -- the plume core emphasised within each image,
    every image weighted alike."""
    Tp = np.pad(T.astype(np.float64), ((0, 0), (1, 1), (1, 1)), mode="edge")
    h, w = T.shape[1:]
    tb = sum(Tp[:, a:a + h, b:b + w] for a in range(3) for b in range(3)) / 9.0
    t2 = tb * tb
    m2 = t2.mean(axis=(1, 2), keepdims=True)
    if kind == "core":
        wgt = (t2 + m2) / np.maximum(2.0 * m2, 1e-30)
        wgt = np.where(m2 > 0, wgt, 1.0)
    else:
        m4 = np.maximum((t2 * t2).mean(axis=(1, 2), keepdims=True), 1e-30)
        wgt = (t2 + 0.1 * m2) / m4
    return (wgt / wgt.mean()).astype(np.float32)


def augment(x, t, rng, w=None):
    """The same random rotation and flip of the image axes of x ([n][h][w] or
    [n][frames][h][w]), t (and the loss weights w)."""
    k = int(rng.integers(4))
    flip = bool(rng.integers(2))

    def tr(a):
        if a is None:
            return None
        a = np.rot90(a, k, axes=(-2, -1))
        return np.ascontiguousarray(a[..., ::-1] if flip else a)
    return (tr(x), tr(t)) if w is None else (tr(x), tr(t), tr(w))


def train(net: UNet, X, T, epochs=20, batch=16, lr=1e-3, lr_min=1e-4, seed=1, log=None, weights=None):
    """Adam with a cosine learning-rate decay; random rotations and flips;
    per-pixel loss weights when given."""
    rng = np.random.default_rng(seed)
    n = len(X)
    steps_per_epoch = max(1, n // batch)
    total = epochs * steps_per_epoch
    hist = []
    t0 = time.perf_counter()
    step = 0
    for ep in range(epochs):
        order = rng.permutation(n)
        losses = []
        for s in range(steps_per_epoch):
            b = order[s * batch:(s + 1) * batch]
            a = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * step / total))
            if weights is None:
                x, t = augment(X[b], T[b], rng)
                losses.append(net.train_step(x, t, a))
            else:
                x, t, w = augment(X[b], T[b], rng, weights[b])
                losses.append(net.train_step(x, t, a, weight=w))
            step += 1
        hist.append(float(np.mean(losses)))
        if log:
            log(f"  epoch {ep + 1}/{epochs}: loss {hist[-1]:.4f}")
    return dict(loss=hist, seconds=time.perf_counter() - t0, steps=step)


def _dihedral(a, k, flip, inverse=False):
    """Rotation by k quarter turns (then a flip of the last axis) of the
    image axes, or its inverse."""
    if not inverse:
        a = np.rot90(a, k, axes=(-2, -1))
        return a[..., ::-1] if flip else a
    a = a[..., ::-1] if flip else a
    return np.rot90(a, -k, axes=(-2, -1))


def reconstruct(net, few, sub_every, scale, batch=16, extra=None, tta=False):
    """Concentration estimate (counts of the full release) from few-particle
    histograms (and the extra planes the network was trained with), and the
    device ms of the forward passes.  net: a UNet or a list of them (their
    outputs averaged).  tta: average each network's output over the 8
    rotations and flips of the input (each mapped back), in the network's
    units."""
    nets = net if isinstance(net, (list, tuple)) else [net]
    x, _ = to_net(few, None, sub_every, scale, extra)
    out = np.zeros((len(x),) + x.shape[-2:], np.float64)
    ms = 0.0
    views = [(k, f) for k in range(4) for f in (False, True)] if tta else [(0, False)]
    for nt in nets:
        for k, f in views:
            xv = np.ascontiguousarray(_dihedral(x, k, f))
            for b in range(0, len(x), batch):
                y, m = nt.forward(xv[b:b + batch])
                out[b:b + batch] += _dihedral(y, k, f, inverse=True)
                ms += m
    out /= len(views) * len(nets)
    return (np.maximum(out, 0.0) * scale) ** 2, ms


def rel_l2(est, ref):
    """Per-image relative L2 error."""
    d = (est - ref).reshape(len(ref), -1)
    r = ref.reshape(len(ref), -1).astype(np.float64)
    return np.linalg.norm(d, axis=1) / np.maximum(np.linalg.norm(r, axis=1), 1e-30)


def windowed(net, few, sub, scale, window, extra=None, tta=False):
    """Reconstruction of whole-surface histograms few [n][H][W] (or
    [n][frames][H][W]): the network runs on the window x window crop around
    the last frame's centre of mass (the plume; the release point is not
    needed) and plain gridding is the fallback outside it.  Returns the
    estimates [n][H][W] and the device ms of the forward passes."""
    last = few[:, 0] if few.ndim == 4 else few
    n, H, W = last.shape
    jj, ii = np.arange(H, dtype=np.float64), np.arange(W, dtype=np.float64)
    est = sub * last.astype(np.float64)
    org = []
    for q in range(n):
        f = last[q].astype(np.float64)
        tot = f.sum()
        cj = (f.sum(1) @ jj) / tot if tot > 0 else H / 2
        ci = (f.sum(0) @ ii) / tot if tot > 0 else W / 2
        org.append((int(np.clip(round(cj) - window // 2, 0, H - window)),
                    int(np.clip(round(ci) - window // 2, 0, W - window))))
    crop = lambda a, q: a[..., org[q][0]:org[q][0] + window, org[q][1]:org[q][1] + window]   # noqa: E731
    fc = np.stack([crop(few[q], q) for q in range(n)])
    xc = None if extra is None else np.stack([crop(extra[q], q) for q in range(n)])
    rec, ms = reconstruct(net, fc, sub, scale, extra=xc, tta=tta)
    for q in range(n):
        r0, c0 = org[q]
        est[q, r0:r0 + window, c0:c0 + window] = rec[q]
    return est, ms


def evaluate(net, data, scale, extras=(), tta=False, window=256):
    """Held-out errors of plain gridding and of the network against the
    N-particle histograms, the targets' own noise, and the estimated errors
    against the noise-free field.  Images whose target is empty (every
    particle left the surface layer or the crop) carry no information and
    are left out (their number is reported).  Whole-surface images are
    reconstructed in a window x window plume window (windowed)."""
    keep = data["all"].reshape(len(data["all"]), -1).sum(1) > 0
    few, allc, sub = data["few"][keep], data["all"][keep].astype(np.float64), data["sub_every"]
    extra = extra_planes(data, extras)
    grid = sub * (few[:, 0] if few.ndim == 4 else few).astype(np.float64)      # the last hour's histogram
    if data.get("full"):
        rec, ms = windowed(net, few, sub, scale, window, extra=None if extra is None else extra[keep], tta=tta)
    else:
        rec, ms = reconstruct(net, few, sub, scale, extra=None if extra is None else extra[keep], tta=tta)
    e_grid, e_net = rel_l2(grid, allc), rel_l2(rec, allc)
    e_t = np.sqrt(allc.reshape(len(allc), -1).sum(1)) / np.linalg.norm(allc.reshape(len(allc), -1), axis=1)
    truth = lambda e: np.sqrt(np.maximum(e ** 2 - e_t ** 2, 0.0))   # noqa: E731
    return dict(e_grid=float(e_grid.mean()), e_net=float(e_net.mean()), e_target=float(e_t.mean()),
                e_grid_truth=float(truth(e_grid).mean()), e_net_truth=float(truth(e_net).mean()),
                net_better_fraction=float((e_net < e_grid).mean()), forward_ms_per_image=ms / len(few),
                images=len(few), empty_targets=int((~keep).sum()))


def train_flops_per_image(h: int, w: int) -> float:
    """Multiply-add FLOPs of one training step per image (forward + both
    backward products of every 3x3 conv): 3 x sum 2 H W 9 cin cout."""
    from .types import UNET_SPEC
    f, level = 0.0, [0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0]
    for (co, c1, c2), lv in zip(UNET_SPEC, level):
        f += 2.0 * (h >> lv) * (w >> lv) * 9 * (c1 + c2) * co
    return 3.0 * f


def run(ctx: Context, W: ForcingWindow, n_train: int = 512, n_test: int = 128, epochs: int = 40,
        n_target: int = 1_000_000, sub_every: int = 20, hours: int = 24, crop: int = 128, batch: int = 16,
        scale: float = 16.0, seed: int = 21, frames: int = 1, extras=("kh", "mask"), loss_weight: str = "none",
        tta: bool = True, members: int = 1, log=None, pairs=None, eval_full: bool = False):
    """The whole learned-reconstruction experiment: pairs, training, held-out
    errors, and the comparison with plain gridding at equal GPU time (the
    move of n = n_target / sub_every particles per scenario, batched 256 at a
    time, plus one forward pass per image).  frames > 1: the network also
    sees the few-particle histograms of the frames - 1 hours before; extras:
    the extra input planes (EXTRAS); loss_weight: "none", "count" or "core"
    (loss_weights); tta: average over the 8 rotations and flips at
    inference (the single-pass errors are reported beside); members:
    networks trained from different initialisations whose outputs are
    averaged.  pairs: (train, test) data from an earlier call with
    the same sizes, reused instead of generated.  eval_full: the held-out
    images are the whole eta x xi surface layer, not crops."""
    if pairs is None:
        tr = generate_pairs(ctx, W, n_train, n_target, sub_every, hours, crop, seed, frames=frames, log=log)
        te = generate_pairs(ctx, W, n_test, n_target, sub_every, hours, None if eval_full else crop, seed + 1,
                            frames=frames, log=log)
    else:
        tr, te = pairs
    X, T = to_net(tr["few"], tr["all"], sub_every, scale, extra_planes(tr, extras))
    if loss_weight not in ("none", "count", "core"):
        raise ValueError(f"loss_weight {loss_weight!r}: 'none', 'count' or 'core'")
    nets, hs = [], []
    for m in range(members):
        net = UNet(ctx_device(ctx), seed=3 + m, channels=frames + len(extras))
        hs.append(train(net, X, T, epochs=epochs, batch=batch, log=log, seed=1 + m,
                        weights=None if loss_weight == "none" else loss_weights(T, loss_weight)))
        nets.append(net)
    h = dict(loss=hs[0]["loss"], seconds=sum(x["seconds"] for x in hs), steps=sum(x["steps"] for x in hs))
    ev = evaluate(nets, te, scale, extras, tta=tta, window=crop)
    single = evaluate(nets[0], te, scale, extras, window=crop) if (tta or members > 1) else None
    for net in nets:
        net.close()
    # GPU time of the few-particle move: 256 scenarios of n particles, batched
    n = n_target // sub_every
    srcs = [EnsSource(k=-0.27, j=float(tr["j"][q % n_train]), i=float(tr["i"][q % n_train]), spread=0.25,
                      time=float(W.times[0]), count=n, scenario=q, sub_scenario=-1, sub_every=0) for q in range(256)]
    ctx.ens_release(srcs, np.full(256, math.sqrt(60.0)), seed=seed + 2)
    cfg = default_config(random=True, sigma=1.0)
    st = ctx.ens_run(cfg, [float(W.times[q]) for q in range(hours + 1)], list(range(hours + 1)), use_graph=True,
                     surface=True)
    move_ms = st.device_ms / 256
    net_ms = ev["forward_ms_per_image"]
    n_equal = n * (move_ms + net_ms) / move_ms              # plain gridding's particles at the same GPU time
    e_grid_equal = ev["e_grid_truth"] * math.sqrt(n / n_equal)
    n_match = n * (ev["e_grid_truth"] / max(ev["e_net_truth"], 1e-12)) ** 2
    flops = train_flops_per_image(crop, crop) * h["steps"] * batch
    return dict(
        pairs=dict(train=n_train, test=n_test, n_target=n_target, n_input=n, hours=hours, crop=crop, frames=frames,
                   held_out_images=(f"whole surface {W.eta}x{W.xi}; the network in the {crop}x{crop} window around "
                                    "the input's centre of mass, plain gridding outside") if eval_full
                   else f"{crop}x{crop} crops",
                   extras=list(extras), loss_weight=loss_weight, tta=tta, members=members,
                   seconds=tr["seconds"] + te["seconds"], particle_steps=int(tr["particle_steps"] + te["particle_steps"])),
        training=dict(epochs=epochs, steps=h["steps"], seconds=h["seconds"], loss_first=h["loss"][0],
                      loss_last=h["loss"][-1], tflops=flops / h["seconds"] / 1e12),
        held_out=ev,
        held_out_single_pass=single,
        equal_gpu_ms=dict(move_ms_per_scenario=move_ms, net_ms_per_image=net_ms, plain_particles=n_equal,
                          plain_e_truth=e_grid_equal, learned_e_truth=ev["e_net_truth"],
                          plain_particles_to_match_learned=n_match),
    )


def ctx_device(ctx: Context) -> int:
    return ctx.device


def time_training(steps: int, batch: int = 16, size: int = 128, seed: int = 1) -> dict:
    """Times `steps` training steps and one forward pass of the network on a
    fixed random batch (the training kernels' throughput, without data)."""
    rng = np.random.default_rng(seed)
    net = UNet(0, seed=seed)
    x = rng.uniform(0.0, 2.0, (batch, size, size)).astype(np.float32)
    t = (x + rng.normal(0.0, 0.1, x.shape)).astype(np.float32)
    for _ in range(5):
        net.train_step(x, t, 1e-3)
    t0 = time.perf_counter()
    for _ in range(steps):
        net.train_step(x, t, 1e-3)
    dt = time.perf_counter() - t0
    _, fwd_ms = net.forward(x)
    net.close()
    f = train_flops_per_image(size, size) * batch
    return dict(steps=steps, batch=batch, size=size, ms_per_step=1e3 * dt / steps, forward_ms=fwd_ms,
                train_tflops=f * steps / dt / 1e12, forward_tflops=f / 3 / (fwd_ms * 1e-3) / 1e12)


if __name__ == "__main__":
    import argparse
    import json
    import sys

    from .synthetic import full_window

    ap = argparse.ArgumentParser(description="train and evaluate the learned concentration reconstruction")
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--target", type=int, default=1_000_000)
    ap.add_argument("--sub-every", type=int, default=20)
    ap.add_argument("--hours", type=int, default=24, help="hours each release is moved")
    ap.add_argument("--crop", type=int, default=128, help="training crop size (cells)")
    ap.add_argument("--eval-full", action="store_true",
                    help="held-out images are the whole 1024 x 768 surface layer instead of crops")
    ap.add_argument("--frames", type=int, default=1, help="hourly few-particle frames the network sees")
    ap.add_argument("--loss-weight", default="none", metavar="SETS",
                    help="loss weighting, none, count or core; several separated by ';' train "
                         "one network each on the same pairs")
    ap.add_argument("--members", type=int, default=1, help="networks averaged (different initialisations)")
    ap.add_argument("--no-tta", dest="tta", action="store_false",
                    help="one forward pass instead of the average over the 8 rotations and flips")
    ap.add_argument("--extras", default="kh,mask", metavar="SETS",
                    help="extra input planes, comma-separated (kh, mask); several sets separated by ';' "
                         "train one network each on the same pairs ('none' = no extras)")
    ap.add_argument("--time-training", type=int, default=0, metavar="STEPS",
                    help="only time STEPS training steps on a random batch of 16 128x128 images")
    a = ap.parse_args()
    if a.time_training:
        print(json.dumps(time_training(a.time_training)), flush=True)
        sys.exit(0)
    say = lambda m: print(m, file=sys.stderr, flush=True)   # noqa: E731
    W = full_window(T=25)
    c = Context(0)
    c.set_grid(W.lonRad, W.latRad, W.h, W.depthIntervals, W.s_rho, mask=W.mask, n_slots=W.T, store="bf16")
    for t in range(W.T):
        c.upload_record(t, *W.record(t))
    sets = [tuple(e for e in grp.split(",") if e and e != "none") for grp in (a.extras or "none").split(";")]
    weights = [x for x in a.loss_weight.split(";") if x]
    combos = [(st, lw) for st in sets for lw in weights]
    if len(combos) == 1:
        print(json.dumps(run(c, W, a.train, a.test, a.epochs, a.target, a.sub_every, hours=a.hours, crop=a.crop,
                             frames=a.frames, extras=combos[0][0], loss_weight=combos[0][1], tta=a.tta,
                             members=a.members, log=say, eval_full=a.eval_full)), flush=True)
    else:
        pairs = (generate_pairs(c, W, a.train, a.target, a.sub_every, seed=21, frames=a.frames, log=say),
                 generate_pairs(c, W, a.test, a.target, a.sub_every, seed=22, frames=a.frames, log=say))
        out = {}
        for st, lw in combos:
            name = f"{','.join(st) or 'none'}/{lw}"
            say(f"== extras / loss weight: {name}")
            out[name] = run(c, W, a.train, a.test, a.epochs, a.target, a.sub_every, frames=a.frames, extras=st,
                            loss_weight=lw, tta=a.tta, members=a.members, log=say, pairs=pairs)
        print(json.dumps({"sets": out}), flush=True)

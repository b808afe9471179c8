"""GPU tests of the LDS-staged forcing of the context's mixed-precision move
(wc_set_lds): every workgroup copies the interval pack of the box its
particles occupy into LDS before each chunk of substeps and samples there
(global memory outside the box).  The values read are the same and a
particle resumes exactly at a chunk boundary, so every particle must end
bit-identical to the plain move's, for any chunk, halo and cell-tile order,
through compaction, with both integrators, both tracking directions and
diffusion on and off; the step count must match too."""
import numpy as np
import pytest

import wacommplusplus_amd as wc
from runner import load_grid
from wacommplusplus_amd.synthetic import full_window, release

pytestmark = pytest.mark.gpu

FIELDS = ("id", "k", "j", "i", "health", "age", "time")


@pytest.fixture(scope="module")
def full():
    return full_window(T=5)


def run(ctx, W, cfg, p, store, lds, tile, n=4, backward=False):
    load_grid(ctx, W, W.T, store)
    for t in range(W.T):
        ctx.upload_record(t, *W.record(t))
    ctx.set_sort(tile, 1)
    ctx.set_lds(*lds)
    ctx.set_particles(p)
    ctx.reset_step_count()
    order = range(n - 1, -1, -1) if backward else range(n)
    for s in order:
        a, b = (s + 1, s) if backward else (s, s + 1)
        ctx.move_interval(cfg, float(W.times[a]), float(W.times[b]), a, b, path="fast")
        ctx.remove_dead()
    out, steps = ctx.get_particles(), ctx.step_count()
    ctx.set_sort(0)
    ctx.set_lds(0)
    return out, steps


@pytest.mark.parametrize("store", ["bf16", "f16", "f32"])
def test_lds_move_equals_plain_move(full, store):
    p = release(full, 4_000_000, 50, spread=2.0)[::20].copy()            # 200k particles of the full release
    cfg = wc.default_config(random=True, sigma=wc.sigma_from_kh(1.0, 30.0))
    with wc.Context(0) as ctx:
        for tile in (0, 1):
            ref, ref_steps = run(ctx, full, cfg, p, store, (0, 0), tile)
            for lds in ((1, 0), (10, 2), (30, 2), (120, 8)):
                got, steps = run(ctx, full, cfg, p, store, lds, tile)
                print(f"{store} tile {tile} lds {lds}: {len(got)} particles, {steps} steps", flush=True)
                assert steps == ref_steps, (tile, lds)
                for f in FIELDS:
                    assert np.array_equal(got[f], ref[f], equal_nan=f != "id"), (tile, lds, f)


@pytest.mark.parametrize("integrator,random,backward", [(wc.EULER, True, False), (wc.MIDPOINT, False, False),
                                                        (wc.MIDPOINT, True, True)])
def test_lds_move_equals_plain_move_variants(full, integrator, random, backward):
    p = release(full, 4_000_000, 50, spread=2.0)[3::40].copy()
    cfg = wc.default_config(random=random, sigma=wc.sigma_from_kh(1.0, 30.0), integrator=integrator,
                            horizontalClosure=wc.KILL, trackingDirection=-1 if backward else 1,
                            backwardDiffusion=1 if backward else 0)
    with wc.Context(0) as ctx:
        ref, ref_steps = run(ctx, full, cfg, p, "bf16", (0, 0), 1, backward=backward)
        got, steps = run(ctx, full, cfg, p, "bf16", (15, 1), 1, backward=backward)
        assert steps == ref_steps
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f], equal_nan=f != "id"), f


def test_lds_setting_is_checked_and_kept():
    with wc.Context(0) as ctx:
        assert ctx.get_lds() == (0, 0)
        ctx.set_lds(30, 2)
        assert ctx.get_lds() == (30, 2)
        for bad in ((-1, 0), (1025, 0), (10, 17)):
            with pytest.raises(wc.WcError):
                ctx.set_lds(*bad)
        assert ctx.get_lds() == (30, 2)

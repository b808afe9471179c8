"""Writes tests/golden/ref_vectors.npz: inputs and outputs of the reference's
own helpers, compiled from its sources by `make -C oracle ref` into
oracle/_ref/libwcref.so (NumericalHelpers.hpp, SourceEmission.hpp and the
inline helpers of DriftModel.hpp).  The file is data -- seeded inputs and the
values the reference computed for them -- so the oracle can be checked
against the reference wherever the tests run, the reference sources absent.

    make -C oracle ref && python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle as po  # noqa: E402


def main():
    R = po.ref_lib()
    if R is None or not hasattr(R, "ref_jibe_step_probability"):
        sys.exit("oracle/_ref/libwcref.so missing or stale: run `make -C oracle ref` first")
    rng = np.random.default_rng(20261203)
    out = {}
    # keyed RNG: uniform and normal on random keys, mix on random words
    n = 1000
    key = np.stack([rng.integers(0, 2**63, n, dtype=np.uint64) * 2 + rng.integers(0, 2, n, dtype=np.uint64),
                    rng.integers(0, 2**63, n, dtype=np.uint64),
                    rng.integers(-2**40, 2**40, n).astype(np.int64).view(np.uint64),
                    rng.integers(0, 1 << 20, n, dtype=np.uint64),
                    rng.integers(0, 8, n, dtype=np.uint64)], axis=1)
    key[:3, 2] = np.array([-2**63, -2**63 + 1, -2**63 + 2], dtype=np.int64).view(np.uint64)   # reserved intervals
    out["rng_key"] = key
    out["rng_uniform"] = np.array([R.ref_uniform(int(a), int(b), int(np.int64(c.view(np.int64))), int(d), int(e))
                                   for a, b, c, d, e in key])
    out["rng_normal"] = np.array([R.ref_normal(int(a), int(b), int(np.int64(c.view(np.int64))), int(d), int(e))
                                  for a, b, c, d, e in key])
    words = rng.integers(0, 2**63, 500, dtype=np.uint64) * 2 + rng.integers(0, 2, 500, dtype=np.uint64)
    out["mix_in"] = words
    out["mix_out"] = np.array([R.ref_mix(int(w)) for w in words], dtype=np.uint64)
    # numerical helpers
    m = 1000
    se = rng.uniform(-1e4, 1e4, (m, 3))
    se[::7, 2] = se[::7, 0] + rng.uniform(0, 60, len(se[::7]))          # checkpoints inside the interval
    out["restart_in"] = se
    out["restart_out"] = np.array([R.ref_restart_elapsed(*r) for r in se])
    old = rng.uniform(0, 50, m)
    cand = old + rng.uniform(-2, 2, m)
    out["reflect_cell_in"] = np.stack([old, cand], axis=1)
    out["reflect_cell_out"] = np.array([R.ref_reflect_cell(o, int(np.floor(o)), c, int(np.floor(c)))
                                        for o, c in zip(old, cand)])
    x, mx = rng.uniform(-300, 300, m), rng.integers(1, 120, m).astype(np.float64)
    out["reflect_domain_in"] = np.stack([x, mx], axis=1)
    out["reflect_domain_out"] = np.array([R.ref_reflect_domain(a, b) for a, b in zip(x, mx)])
    # emission schedules (ragged: times concatenated, counts alongside; -1 = rejected)
    sched, counts, times = [], [], []
    for _ in range(200):
        mode = int(rng.integers(0, 3))
        value = float(rng.integers(0, 50)) if mode else float(rng.uniform(0, 300))
        s0 = float(rng.uniform(0, 20000))
        s1 = -1.0 if rng.random() < .3 else s0 + float(rng.uniform(0, 20000))
        i0 = float(rng.integers(0, 10)) * 3600.0
        sched.append((mode, value, s0, s1, i0, i0 + 3600.0))
        t = po.emission_times(mode, value, s0, s1, i0, i0 + 3600.0, use_ref=True)
        if isinstance(t, int):
            counts.append(-1)
        else:
            counts.append(len(t))
            times.extend(t)
    out["emission_in"] = np.array(sched)
    out["emission_count"] = np.array(counts, dtype=np.int64)
    out["emission_times"] = np.array(times)
    # source position offsets (SourceEmission.hpp:20-23)
    sp = np.stack([rng.integers(0, 2**63, 300, dtype=np.uint64), rng.integers(0, 1 << 30, 300, dtype=np.uint64),
                   rng.integers(0, 3, 300, dtype=np.uint64)], axis=1)
    out["source_offset_in"] = sp
    out["source_offset_out"] = np.array([R.ref_source_position_offset(int(a), int(b), int(c)) for a, b, c in sp])
    # drift helpers
    hp = np.array([(h, dt) for h in (-0.5, 0.0, 1e-9, 0.05, 0.3, 0.5, 0.999, 1.0, 2.0)
                   for dt in (-1.0, 0.0, 1e-3, 5.0, 30.0, 3600.0, 7200.5)])
    out["jibe_in"] = hp
    out["jibe_out"] = np.array([R.ref_jibe_step_probability(h, dt) for h, dt in hp])
    out["side_seed"] = rng.integers(0, 2**63, 500, dtype=np.uint64)
    out["side_id"] = rng.integers(0, 2**63, 500, dtype=np.uint64)
    out["side_prob"] = rng.uniform(-0.1, 1.1, 500)
    out["side_out"] = np.array([R.ref_sample_drift_side(int(a), int(b), float(c)) for a, b, c in
                                zip(out["side_seed"], out["side_id"], out["side_prob"])], dtype=np.int64)
    out["jibe_component"] = np.array([R.ref_jibe_component()], dtype=np.uint64)
    # rank partition
    pt = np.array([(t, w, r) for t in (0, 1, 11, 4_000_003, 2**40 + 5) for w in (1, 2, 3, 8, 13) for r in range(w)],
                  dtype=np.uint64)
    out["partition_in"] = pt
    out["partition_out"] = np.array([(R.ref_partition_offset(int(t), int(w), int(r)),
                                      R.ref_partition_count(int(t), int(w), int(r))) for t, w, r in pt], dtype=np.uint64)
    out["scheme"] = np.frombuffer(R.ref_scheme(), dtype=np.uint8)
    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()

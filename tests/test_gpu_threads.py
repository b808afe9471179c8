"""Host threading of the context API: one context per caller, calls on one
context serialised by its caller, any host thread may drive it (the CUDA
backend's caller drives its per-GPU streams from the master thread,
Wacomm.cpp:542-581; the library holds no global mutable state).

Two contexts moved concurrently from two threads (ctypes releases the GIL
during every library call), and a context created on one thread and driven
from another, give bit-identical results to the same runs one after another.
"""
import threading

import numpy as np
import pytest

import wacommplusplus_amd as wc
from runner import run_device
from test_gpu_parity import random_window

pytestmark = pytest.mark.gpu


def population(W, n, seed):
    rng = np.random.default_rng(seed)
    p = wc.make_particles(n)
    p["k"] = rng.uniform(-2.5, -0.1, n)
    p["j"] = rng.uniform(1.0, W.eta - 2.0, n)
    p["i"] = rng.uniform(1.0, W.xi - 2.0, n)
    p["id"] = rng.integers(0, 2**63, n, dtype=np.uint64)
    return p


def test_two_contexts_on_two_threads():
    W = random_window(2)
    jobs = [(population(W, 50_000, 1), wc.default_config(random=True, sigma=30.0, randomSeed=4), "parity"),
            (population(W, 50_000, 2), wc.default_config(random=True, sigma=30.0, randomSeed=5), "fast")]
    # sequential reference
    seq = []
    for p, cfg, path in jobs:
        with wc.Context(0) as c:
            seq.append(run_device(c, W, cfg, p, 3, path=path, store="bf16" if path == "fast" else "f32"))
    # concurrent, each thread with its own context, several rounds
    for _ in range(3):
        out, errors = [None, None], []

        def work(q):
            try:
                p, cfg, path = jobs[q]
                with wc.Context(0) as c:
                    out[q] = run_device(c, W, cfg, p, 3, path=path, store="bf16" if path == "fast" else "f32")
            except Exception as e:          # reported below
                errors.append(e)
        th = [threading.Thread(target=work, args=(q,)) for q in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join(timeout=120)
        assert not errors, errors
        for q in range(2):
            for f in seq[q].dtype.names:
                assert np.array_equal(out[q][f], seq[q][f], equal_nan=f not in ("id", "driftObjectType", "driftSide")), (q, f)


def test_context_driven_from_another_thread():
    W = random_window(3)
    p = population(W, 20_000, 3)
    cfg = wc.default_config(random=True, sigma=30.0)
    with wc.Context(0) as c:
        ref = run_device(c, W, cfg, p, 2)
    c = wc.Context(0)                        # created here ...
    out, errors = [], []

    def work():
        try:
            out.append(run_device(c, W, cfg, p, 2))      # ... driven there
        except Exception as e:
            errors.append(e)
    t = threading.Thread(target=work)
    t.start()
    t.join(timeout=120)
    c.close()
    assert not errors, errors
    for f in ref.dtype.names:
        assert np.array_equal(out[0][f], ref[f], equal_nan=f not in ("id", "driftObjectType", "driftSide")), f

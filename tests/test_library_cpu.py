"""CPU checks of libwacomm_hip.so: it loads, exports every symbol the C header
declares, shares the reference's struct layouts with the oracle, and its host
utilities agree with the oracle and with published vectors.  No kernel is
launched here (this container has no GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle as po
import wacommplusplus_amd as wc
from wacommplusplus_amd.synthetic import full_window, plain_window, depth_to_k

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "wacomm_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(wc_[a-z0-9_]+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    L = wc.load_library()
    declared = header_symbols()
    assert declared == set(wc.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name
    assert L.wc_version().startswith(b"wacomm-hip")


def test_struct_layouts_match_reference_and_oracle():
    assert C.sizeof(wc.ConfigData) == C.sizeof(po.Config)
    for (n1, t1), (n2, t2) in zip(wc.ConfigData._fields_, po.Config._fields_):
        assert n1 == n2 and getattr(wc.ConfigData, n1).offset == getattr(po.Config, n2).offset
    assert wc.PARTICLE_DTYPE == po.PARTICLE_DTYPE and wc.PARTICLE_DTYPE.itemsize == 64


def test_context_without_device_fails_cleanly():
    # no GPU in this container: creating a context must report, not crash
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(wc.WcError) as e:
        wc.Context(0)
    assert e.value.status == 5


def test_metric_table_equals_oracle():
    W = full_window(T=1, eta=64, xi=48)
    dx, dy = wc.cell_metric_table(W.lonRad, W.latRad)
    g = po.Window(W.times, W.lonRad, W.latRad, W.h, W.depthIntervals,
                  *[a[None] for a in W.record(0)])
    for j in range(0, 63, 3):
        for i in range(0, 47, 3):
            assert g.cell_metric(j, i) == (dx[j, i], dy[j, i])


def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32-10 (Salmon et al., SC'11)
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, want in kat:
        assert wc.philox4x32_10(ctr, key).tolist() == want


def test_depth_to_k():
    # ApplicationEndToEnd: 10 m release at the grid of ParticleMoveTest-like sigma gives k=-0.2
    assert abs(depth_to_k(1.0, 50.0, 11) - (-0.2)) < 1e-12
    assert abs(depth_to_k(1.0, 40.0, 12) - (-0.275)) < 1e-12
    assert depth_to_k(0.0, 40.0, 12) == 0.0
    assert depth_to_k(100.0, 40.0, 12) == -10.0


def test_synthetic_full_window_properties():
    W = full_window(T=2, eta=256, xi=192)
    z, u, v, w, a = W.record(0)
    land = W.mask == 0
    assert 0.25 < land.mean() < 0.35
    assert (W.h[land] == 0).all() and (u[:, land] == 0).all() and (w[:, land] == 0).all()
    speed = np.sqrt(u[-1] ** 2 + v[-1] ** 2)[~land]
    assert 0.15 < np.sqrt((speed ** 2).mean()) < 0.45
    assert (w[0] == 0).all()                      # no flow through the bottom


def test_reconstruction_inputs():
    """Network input of the learned reconstruction: histogram frames then the
    extra planes as channels; rotations and flips act on every channel alike."""
    from wacommplusplus_amd.reconstruct import augment, extra_planes, to_net
    rng = np.random.default_rng(0)
    n, h, w = 3, 16, 16
    data = dict(all=rng.integers(0, 50, (n, h, w)).astype(np.uint32), kh=np.array([0.1, 1.0, 10.0]),
                mask=(rng.random((n, h, w)) < 0.7).astype(np.uint8))
    few = rng.integers(0, 5, (n, h, w)).astype(np.uint32)
    ex = extra_planes(data, ("kh", "mask"))
    assert ex.shape == (n, 2, h, w) and ex.dtype == np.float32
    assert np.allclose(ex[:, 0, 3, 5], [-1.0, 0.0, 1.0])
    assert np.array_equal(ex[:, 1], data["mask"])
    x, t = to_net(few, data["all"], 20, 16.0, ex)
    assert x.shape == (n, 3, h, w) and t.shape == (n, h, w)
    assert np.allclose(x[:, 0], np.sqrt(20.0 * few) / 16.0)
    assert extra_planes(data, ()) is None
    x1, _ = to_net(few, None, 20, 16.0)
    assert x1.shape == (n, h, w)
    for seed in range(8):
        xa, ta = augment(x, t, np.random.default_rng(seed))
        # one rotation/flip, the same for every channel and the target
        transforms = [lambda a, q=q, f=f: (np.rot90(a, q, axes=(-2, -1))[..., ::-1] if f
                                          else np.rot90(a, q, axes=(-2, -1))) for q in range(4) for f in (0, 1)]
        hits = [T for T in transforms if np.array_equal(T(x), xa) and np.array_equal(T(t), ta)]
        assert hits, seed
    with pytest.raises(ValueError):
        extra_planes(data, ("wind",))


def test_reconstruction_loss_weights():
    """The count-aligned loss weights: positive, mean one, growing with the
    (box-filtered) target, and moved by augmentation with the images."""
    from wacommplusplus_amd.reconstruct import augment, loss_weights
    rng = np.random.default_rng(3)
    T = np.zeros((4, 16, 16), np.float32)
    T[:, 6:10, 6:10] = rng.uniform(1.0, 2.0, (4, 4, 4))
    w = loss_weights(T)
    assert w.shape == T.shape and w.dtype == np.float32
    assert np.all(w > 0) and abs(float(w.mean()) - 1.0) < 1e-5
    assert np.all(w[:, 8, 8] > 10 * w[:, 0, 0])
    x = rng.uniform(0, 1, (4, 2, 16, 16)).astype(np.float32)
    xa, ta, wa = augment(x, T, np.random.default_rng(1), w)
    for q in range(4):
        for f in (0, 1):
            R = lambda a: np.rot90(a, q, axes=(-2, -1))[..., ::-1] if f else np.rot90(a, q, axes=(-2, -1))  # noqa: E731
            if np.array_equal(R(T), ta):
                assert np.array_equal(R(w), wa) and np.array_equal(R(x), xa)
                return
    raise AssertionError("augmentation is not one rotation/flip")


def test_reconstruction_tta_views_invert():
    """Each of the 8 rotation/flip views used for test-time averaging maps
    back to the original image, and the 8 views are distinct."""
    from wacommplusplus_amd.reconstruct import _dihedral
    a = np.arange(2 * 3 * 8 * 8, dtype=np.float32).reshape(2, 3, 8, 8)
    views = []
    for k in range(4):
        for f in (False, True):
            v = _dihedral(a, k, f)
            assert np.array_equal(_dihedral(v, k, f, inverse=True), a)
            views.append(v.tobytes())
    assert len(set(views)) == 8


def test_dropin_rejects_populations_beyond_32bit_indexing():
    """Populations above 2^31 - 256 would overflow the kernels' 32-bit
    indexing: the drop-in entry points refuse them before any device call
    (so this runs without a GPU)."""
    L = wc.load_library()
    frame = np.zeros(8, np.float32)
    big = 0x7FFFFF00 + 1
    assert L.wc_hip_concentration(None, big, 2, 2, 2, frame.ctypes.data_as(C.c_void_p), None) == wc.WC_ERR_INVALID
    dummy = np.zeros(64, np.float64)
    d = dummy.ctypes.data_as(C.c_void_p)
    assert L.wc_hip_move_particles(d, d, 0, 2, 3, 2, 2, 2, *([d] * 6), *([d] * 9), big, 1, 1, None) == wc.WC_ERR_INVALID

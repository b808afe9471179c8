"""Small forcing windows used by the tests.

tiny_window() is the 2x2 grid of the reference's own move tests
(tests/ParticleMoveTest.cpp:10-23 and tests/CudaParticleParityTest.cu:35-70):
one-degree cells at the equator, h=100 m, s_w=3 W levels (s_rho=2), zeta=0.
"""
import numpy as np

from pyoracle import Window

DEG = 0.017453292519943295


def tiny_window(times=(0.0, 65.0), u=None, v=None, w=None, akt=None, zeta=None):
    T = len(times)
    eta = xi = 2
    s_rho, s_w = 2, 3
    jj, ii = np.meshgrid(np.arange(eta), np.arange(xi), indexing="ij")
    lon = ii * DEG
    lat = jj * DEG
    h = np.full((eta, xi), 100.0)
    depth = np.array([0.5, 0.5, 0.5])
    z = np.zeros((T, eta, xi), np.float32) if zeta is None else zeta
    U = np.zeros((T, s_rho, eta, xi), np.float32) if u is None else u
    V = np.zeros((T, s_rho, eta, xi), np.float32) if v is None else v
    W = np.zeros((T, s_w, eta, xi), np.float32) if w is None else w
    A = np.zeros((T, s_w, eta, xi), np.float32) if akt is None else akt
    return Window(np.array(times, float), lon, lat, h, depth, z, U, V, W, A, mask=np.ones((eta, xi)))


def haversine_1deg():
    """xdist of tests/ParticleMoveTest.cpp:36-37 (one degree of longitude at the equator)."""
    t = np.sin(0.5 * DEG) ** 2
    return 2.0 * np.arctan2(np.sqrt(t), np.sqrt(1.0 - t)) * 6371000.0


def rect_window(eta, xi, s_rho, spacing_m, lat0_deg, T, u, v, w=0.0, akt=0.0, h=50.0, dt=3600.0):
    """Rectilinear lon/lat grid at `spacing_m` around lat0 with uniform fields
    (the plain CPU configuration: 120x100x10 at 200 m, u=0.2, v=0.1)."""
    R = 6371000.0
    s_w = s_rho + 1
    dlat = spacing_m / R
    dlon = spacing_m / (R * np.cos(np.deg2rad(lat0_deg)))
    jj, ii = np.meshgrid(np.arange(eta), np.arange(xi), indexing="ij")
    lat = np.deg2rad(lat0_deg) + jj * dlat
    lon = np.deg2rad(14.0) + ii * dlon
    hh = np.full((eta, xi), h)
    depth = np.full(s_w - 1, 1.0 / (s_w - 1))
    times = np.arange(T) * dt
    z = np.zeros((T, eta, xi), np.float32)
    U = np.full((T, s_rho, eta, xi), u, np.float32)
    V = np.full((T, s_rho, eta, xi), v, np.float32)
    W = np.full((T, s_w, eta, xi), w, np.float32)
    A = np.full((T, s_w, eta, xi), akt, np.float32)
    return Window(times, lon, lat, hh, depth, z, U, V, W, A, mask=np.ones((eta, xi)))

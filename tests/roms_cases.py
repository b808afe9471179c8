"""ROMS inputs for the normalisation tests.

fixture(): the file tests/ROMSAdapterTest.cpp:17-63 writes -- 2 records,
s_rho=2 (s_w = -1, -0.5, 0), rho grid 2x3 at lat 40/41, lon 10/11/12, h=100,
u grid 2x2 holding 2 and 4 in every row, v grid 1x3 holding 6, w=AKt=zeta=0,
all masks 1, angle 0.

random_roms(): a staggered ROMS grid with land (mask_u = mask_rho(j,i) *
mask_rho(j,i+1), mask_v = mask_rho(j,i) * mask_rho(j+1,i) as ROMS derives
them, plus a few independently flipped faces), random angles, random fields,
NetCDF fill values on land faces and zeta fill values on land.
"""
import math

import numpy as np

NC_FILL_FLOAT = np.float32(9.9692099683868690e+36)
NC_FILL_DOUBLE = 9.9692099683868690e+36


def fixture():
    u = np.zeros((2, 2, 2, 2), np.float32)          # [T][s_rho][eta_u][xi_u]
    u[..., 0], u[..., 1] = 2, 4
    return dict(
        s_w=np.array([-1.0, -0.5, 0.0]),
        lon=np.array([[10, 11, 12], [10, 11, 12]], float), lat=np.array([[40, 40, 40], [41, 41, 41]], float),
        h=np.full((2, 3), 100.0), mask_rho=np.ones((2, 3)), mask_u=np.ones((2, 2)), mask_v=np.ones((1, 3)),
        angle=np.zeros((2, 3)), times=np.array([0.0, 3600.0]),
        zeta=np.zeros((2, 2, 3), np.float32), u=u, v=np.full((2, 2, 1, 3), 6, np.float32),
        w=np.zeros((2, 3, 2, 3), np.float32), akt=np.zeros((2, 3, 2, 3), np.float32))


def random_roms(seed=7, T=3, s_rho=5, eta=48, xi=64, land=0.25, spacing_deg=0.01, flips=6):
    rng = np.random.default_rng(seed)
    s_w = s_rho + 1
    mask_rho = (rng.random((eta, xi)) > land).astype(float)
    mask_rho[:, :2] = 1.0                                    # a wet strip where particles start
    mask_u = mask_rho[:, :-1] * mask_rho[:, 1:]
    mask_v = mask_rho[:-1, :] * mask_rho[1:, :]
    for _ in range(flips):                                  # faces ROMS itself would not derive
        mask_u[rng.integers(eta), rng.integers(xi - 1)] = 1.0
        mask_v[rng.integers(eta - 1), rng.integers(xi)] = 0.0
    jj, ii = np.meshgrid(np.arange(eta), np.arange(xi), indexing="ij")
    lon = 14.0 + ii * spacing_deg
    lat = 40.5 + jj * spacing_deg
    h = rng.uniform(20.0, 60.0, (eta, xi))
    angle = rng.uniform(-math.pi, math.pi, (eta, xi))
    s_w_values = np.linspace(-1.0, 0.0, s_w)
    u = rng.normal(0.0, 0.3, (T, s_rho, eta, xi - 1)).astype(np.float32)
    v = rng.normal(0.0, 0.3, (T, s_rho, eta - 1, xi)).astype(np.float32)
    u[:, :, mask_u == 0] = NC_FILL_FLOAT                    # land faces hold fill values
    v[:, :, mask_v == 0] = NC_FILL_FLOAT
    w = rng.normal(0.0, 1e-4, (T, s_w, eta, xi)).astype(np.float32)
    akt = rng.uniform(0.0, 1e-3, (T, s_w, eta, xi)).astype(np.float32)
    zeta = rng.normal(0.0, 0.2, (T, eta, xi)).astype(np.float32)
    zeta[:, mask_rho == 0] = 1e37
    return dict(s_w=s_w_values, lon=lon, lat=lat, h=h, mask_rho=mask_rho, mask_u=mask_u, mask_v=mask_v,
                angle=angle, times=np.arange(T) * 3600.0, zeta=zeta, u=u, v=v, w=w, akt=akt)


def oracle_grid(po, c, earth_relative=False, missing_u=(), missing_v=(), missing_angle=()):
    return po.RomsGrid(len(c["s_w"]) - 1, c["mask_rho"], c["mask_u"], c["mask_v"], c["angle"],
                       earth_relative=earth_relative, missing_u=missing_u, missing_v=missing_v,
                       missing_angle=missing_angle)


def oracle_record(po, g, c, t):
    return po.roms_normalize(g, c["zeta"][t], c["u"][t], c["v"][t], c["w"][t], c["akt"][t])

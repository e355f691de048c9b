"""Synthetic MAS-like coronal fields for the interpolation workload.

A spherical grid of 300 radii (log-spaced, 1-30 Rs) x 200 colatitudes x
400 longitudes holds the 7 MAS fields (Bp, Bt, Br, Vp, Vt, Vr, rho) in MAS
units as FP32 arrays indexed (phi, theta, r) with r fastest
(src/masInterp.c:185).  The background is a Parker solution (400 km/s,
n = 8.3 cm^-3 and B_r = 2.5 nT at 1 AU, 25-day rotation); a seeded
Gaussian CME shell (1500 km/s, 60 deg wide, density x4) moves out between
the two time levels.
"""
import math

import numpy as np

from . import constants as K

V_UNIT = K.MAS_LENGTH_NORM / K.MAS_TIME_NORM            # cm/s per MAS velocity unit
RHO_UNIT = K.MAS_RHO_NORM                               # g/cm^3
B_UNIT = math.sqrt(4.0 * K.PI * K.MAS_RHO_NORM) * V_UNIT  # G per MAS field unit


def meshes(nr=300, nt=200, nphi=400, rmin=1.0, rmax=30.0):
    r = np.geomspace(rmin, rmax, nr).astype(np.float32)
    t = np.linspace(0.0, math.pi, nt).astype(np.float32)
    p = np.linspace(0.0, 2.0 * math.pi, nphi).astype(np.float32)
    return r, t, p


def level(r, t, p, cme_r, seed=0, v_kms=400.0, n_au=8.3, b_au_nT=2.5, period_days=25.0,
          cme_speed=1500.0, cme_width_deg=60.0, cme_dens=4.0):
    rng = np.random.default_rng(seed)
    th0 = 0.5 * math.pi + rng.uniform(-0.2, 0.2)
    ph0 = math.pi + rng.uniform(-0.5, 0.5)
    P, T, R = np.meshgrid(p.astype(np.float64), t.astype(np.float64), r.astype(np.float64), indexing="ij")
    rau = R * K.RSUN / K.AU
    omega = 2.0 * math.pi / (period_days * 86400.0)
    vr = np.full(R.shape, v_kms * 1e5)
    n = n_au / rau ** 2
    br = (b_au_nT * 1e-5) / rau ** 2
    # Gaussian CME shell: radial width 1.5 Rs, angular width cme_width
    cosang = np.sin(T) * math.sin(th0) * np.cos(P - ph0) + np.cos(T) * math.cos(th0)
    ang = np.arccos(np.clip(cosang, -1.0, 1.0))
    w = np.exp(-0.5 * ((R - cme_r) / 1.5) ** 2) * np.exp(-0.5 * (ang / math.radians(cme_width_deg / 2)) ** 2)
    vr = vr + w * (cme_speed - v_kms) * 1e5
    n = n * (1.0 + (cme_dens - 1.0) * w)
    br = br * (1.0 + 0.5 * w)
    bp = -br * omega * (R * K.RSUN) * np.sin(T) / vr
    fields = {
        "Bp": bp / B_UNIT, "Bt": np.zeros_like(br), "Br": br / B_UNIT,
        "Vp": np.zeros_like(vr), "Vt": np.zeros_like(vr), "Vr": vr / V_UNIT,
        "D": n * K.MP / RHO_UNIT,
    }
    order = ("Bp", "Bt", "Br", "Vp", "Vt", "Vr", "D")
    out = [np.ascontiguousarray(fields[k], dtype=np.float32).ravel() for k in order]
    ms = []
    for _ in range(7):
        ms.extend([r, t, p])
    return out, ms


def levels(seed=0, nr=300, nt=200, nphi=400):
    r, t, p = meshes(nr, nt, nphi)
    return level(r, t, p, cme_r=8.0, seed=seed), level(r, t, p, cme_r=12.0, seed=seed)


def prev_from_nodes(nodes, rscale, v_kms=400.0, tdel_days=1.0 / 64):
    """Previous-step node state for the masWind branch: the position one
    radial step back and the analytic Parker values there (code units)."""
    x, y, z, rm = nodes["x"], nodes["y"], nodes["z"], nodes["rmag"]
    step = (v_kms * 1e5 / K.C) * (tdel_days / K.DAY) / rscale
    scale = np.maximum(rm - step, 1.0) / rm
    rold = rm * scale
    rau = rold * rscale
    br = (2.5e-5 / K.MHD_B_NORM) / rau ** 2
    return {"rold_x": x * scale, "rold_y": y * scale, "rold_z": z * scale, "br": br,
            "btheta": np.zeros_like(br), "bphi": -0.3 * br, "vr": np.full_like(br, v_kms * 1e5 / K.C),
            "vtheta": np.zeros_like(br), "vphi": np.zeros_like(br), "density": 8.3 / rau ** 2}

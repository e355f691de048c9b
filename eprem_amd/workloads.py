"""Synthetic workloads on the cube-shell grid.

Node positions follow what the reference's host loop produces for a steady
radial wind: every global step a node is spawned at the inner boundary
(rmag = 1, i.e. r = rScale) on each stream, rotated by the Sun's azimuth
(spawnNewShell, src/simCore.c:629-687), and every older node moves out
radially by V*tDel (moveNodes/delrFlow, src/simCore.c:736-785,
src/flow.c:169-178).  Node k of a stream is therefore at
rmag = 1 + k*V*tDel/rScale and azimuth phi0 - k*omega*tDel: an Archimedean
(Parker) spiral.  The MHD values are the reference's analytic Parker model
(mhdB, mhdV, mhdDensity, src/flow.c:241-403) and the derivative terms are
formed as in updateMhd (src/flow.c:131-149).

Seeded perturbations (radial speed, density, field strength, node jitter)
make every operator branch reachable: compression and rarefaction, both
signs of the parallel velocity change, thinning of close nodes.
"""
import math

import numpy as np

from . import constants as K
from .abi import EpDims


def cube_directions(rows, cols):
    """Unit vectors of the 6*rows*cols stream footpoints, stream index =
    col + cols*(row + rows*face) (initCubeCoords + initSphereCoords,
    src/cubeShellInit.c:529-822; the coordinate walk accumulates cell
    widths the same way)."""
    out = []
    cw, chh = 1.0 / cols, 1.0 / rows
    for face in range(6):
        for row in range(rows):
            for col in range(cols):
                u, u2, w = -0.5, 0.5, 0.5
                u += 0.5 * cw
                u2 -= 0.5 * cw
                w -= 0.5 * chh
                for _ in range(col):
                    u += cw
                    u2 -= cw
                for _ in range(row):
                    w -= chh
                x, y, z = {0: (u, w, 0.5), 1: (u2, w, -0.5), 2: (0.5, u, w), 3: (-0.5, u2, w),
                           4: (u2, 0.5, w), 5: (u, -0.5, w)}[face]
                r = math.sqrt(x * x + y * y + z * z)
                x, y, z = x / r, y / r, z / r
                out.append((math.atan2(y, x), math.acos(z)))
    return out


def parker_nodes(rows, cols, n_nodes, tdel_days=1.0 / 64, flow_cms=400.0e5, density_au=8.3,
                 b_au=2.5e-5, omega=None, rscale=0.005, seed=None, jitter=0.0, perturb=0.0,
                 thin_frac=0.0, smooth=0, cme=None):
    """Node SoA arrays [S, N] for a Parker-spiral snapshot.

    perturb > 0 multiplies V, n and B by seeded smooth random factors
    (relative amplitude `perturb`) that differ between the current and the
    previous time level; jitter > 0 moves nodes by a random fraction of the
    spacing; thin_frac > 0 pulls that fraction of node pairs to within the
    thinning distance dsh_min = 5e-5 AU.  smooth > 0 low-pass filters the
    perturbations over that many nodes (a physical, smooth field along the
    stream).  cme = dict(r_au=..., width_au=..., speed_cms=..., dens=...,
    ang_deg=..., theta0=..., phi0=...) adds a Gaussian CME shell that moved
    from r_au - speed*tDel to r_au during the step, compressing density and
    field by up to `dens` and raising the radial speed inside the shell.
    """
    rng = np.random.default_rng(seed)
    if omega is None:
        omega = (0.004144 / K.MAS_TIME_NORM) * K.TAU
    S = 6 * rows * cols
    tdel = tdel_days / K.DAY
    us = flow_cms / K.C
    bs = b_au / K.MHD_B_NORM
    dirs = cube_directions(rows, cols)
    k = np.arange(n_nodes, dtype=np.float64)
    step = us * tdel / rscale                      # radial move per step (rScale units)
    rm = 1.0 + k * step
    rm_old = np.maximum(rm - step, 1.0)
    out = {f: np.zeros((S, n_nodes)) for f in ("x", "y", "z", "rmag", "bmag", "dlnb", "dlnn", "dupar")}
    for s, (azi0, zen) in enumerate(dirs):
        r = rm.copy()
        if jitter > 0:
            r[1:] += rng.uniform(-jitter, jitter, n_nodes - 1) * step
        azi = azi0 - k * omega * tdel
        x = r * np.sin(zen) * np.cos(azi)
        y = r * np.sin(zen) * np.sin(azi)
        z = r * np.cos(zen) * np.ones_like(r)
        if thin_frac > 0:
            pick = rng.random(n_nodes) < thin_frac
            pick[:2] = False
            for i in np.nonzero(pick)[0]:
                # put node i within 2e-5 AU of node i-1
                d = np.array([x[i] - x[i - 1], y[i] - y[i - 1], z[i] - z[i - 1]])
                d *= (2.0e-5 / rscale) / max(np.linalg.norm(d), 1e-300)
                x[i], y[i], z[i] = x[i - 1] + d[0], y[i - 1] + d[1], z[i - 1] + d[2]
        rr = np.sqrt(x * x + y * y + z * z)
        th = np.arccos(z / rr)
        rau, rau_old = rr * rscale, rm_old * rscale

        def fac(amp):
            if amp <= 0:
                return np.ones(n_nodes), np.ones(n_nodes)
            if smooth > 0:
                ker = np.exp(-0.5 * (np.arange(-3 * smooth, 3 * smooth + 1) / smooth) ** 2)
                ker /= np.sqrt((ker ** 2).sum())
                half = len(ker) // 2
                lp = lambda v: np.convolve(v, ker, mode="full")[half:half + len(v)]
                base = lp(rng.normal(0, amp, n_nodes))
                new = np.exp(np.clip(base + 0.3 * lp(rng.normal(0, amp, n_nodes)), -1, 1))
                return new, np.exp(np.clip(base, -1, 1))
            base = np.cumsum(rng.normal(0, amp, n_nodes)) / np.sqrt(np.arange(1, n_nodes + 1))
            new = np.exp(np.clip(base + rng.normal(0, amp, n_nodes), -1, 1))
            old = np.exp(np.clip(base, -1, 1))
            return new, old

        fv, fv_old = fac(perturb)
        fn, fn_old = fac(perturb)
        fb, fb_old = fac(perturb)
        if cme is not None:
            # angular profile around the CME axis, radial Gaussian shell
            th0, ph0 = cme.get("theta0", 0.5 * math.pi), cme.get("phi0", math.pi)
            cosang = np.sin(th) * math.sin(th0) * np.cos(np.arctan2(y, x) - ph0) + np.cos(th) * math.cos(th0)
            ang = np.arccos(np.clip(cosang, -1.0, 1.0))
            wang = np.exp(-0.5 * (ang / math.radians(cme.get("ang_deg", 60.0) / 2)) ** 2)
            rc, wr = cme["r_au"], cme["width_au"]
            rc_old = rc - cme["speed_cms"] / K.C * tdel * 1.0
            bump = wang * np.exp(-0.5 * ((rau - rc) / wr) ** 2)
            bump_old = wang * np.exp(-0.5 * ((rau_old - rc_old) / wr) ** 2)
            dv = (cme["speed_cms"] - flow_cms) / flow_cms
            fv, fv_old = fv * (1 + dv * bump), fv_old * (1 + dv * bump_old)
            fn, fn_old = fn * (1 + (cme["dens"] - 1) * bump), fn_old * (1 + (cme["dens"] - 1) * bump_old)
            fb, fb_old = fb * (1 + (cme["dens"] - 1) * bump), fb_old * (1 + (cme["dens"] - 1) * bump_old)
        vr, vr_old = us * fv, us * fv_old
        dens, dens_old = density_au / rau ** 2 * fn, density_au / rau_old ** 2 * fn_old
        br = bs / rau ** 2 * fb
        bp = -1.0 * rau * br * (omega / (vr + K.VERYSMALL)) * np.sin(th)
        bmag = np.sqrt(br * br + bp * bp)
        br_o = bs / rau_old ** 2 * fb_old
        bp_o = -1.0 * rau_old * br_o * (omega / (vr_old + K.VERYSMALL)) * np.sin(th)
        bmag_old = np.sqrt(br_o * br_o + bp_o * bp_o)
        dupar = ((vr - vr_old) * br) / bmag
        dlnb = np.log(bmag / bmag_old)
        dlnn = np.log(dens / dens_old)
        dupar[0] = dlnb[0] = dlnn[0] = 0.0
        out["x"][s], out["y"][s], out["z"][s], out["rmag"][s] = x, y, z, rr
        out["bmag"][s], out["dlnb"][s], out["dlnn"][s], out["dupar"][s] = bmag, dlnb, dlnn, dupar
    return out


def seeded_f(p, grids, nodes, noise=0.0, seed=None):
    """f[S, N, E, M] = sepSeedFunction(E, rmag) (initEnergeticParticles)
    times seeded log-normal noise (anisotropic when noise > 0)."""
    rng = np.random.default_rng(seed)
    eg = grids["egrid"]
    rm = nodes["rmag"]
    efac = K.MEV / (K.MP * K.C * K.C)
    normJ0 = p.boundaryFunctAmplitude * (K.MP * K.C) / (K.MHD_DENSITY_NORM * K.MEV)
    normR = p.boundaryFunctR0 / p.rScale
    radial = (rm / normR) ** (-p.boundaryFunctBeta)
    plt = (eg / (1.0 * efac)) ** (-p.boundaryFunctGamma)
    ext = np.exp(-(eg / (p.boundaryFunctEcutoff * efac)))
    f = normJ0 * radial[:, :, None] * plt[None, None, :] * ext[None, None, :]
    f = f / p.boundaryFunctXi / (2.0 * eg[None, None, :])
    f = np.maximum(f, 2.2250738585072014e-308)
    M = len(grids["mugrid"])
    f = np.repeat(f[..., None], M, axis=3)
    if noise > 0:
        f = f * np.exp(rng.normal(0, noise, f.shape))
    return np.ascontiguousarray(f)


def dims_of(nodes, E, M):
    S, N = nodes["x"].shape
    return EpDims(S, N, E, M)


def power_law_f(grids, nodes, M, seed=None, f_max=1.0e5, index=(-5.0, -2.0), radial_power=-2.0):
    """f[S, N, E, M] for the mixed-precision workload: per stream a seeded
    power law E^g with g uniform in `index`, times (r/r_inner)^radial_power,
    scaled to f_max at the lowest energy of the inner node, isotropic.  With
    the default indices and 10 keV - 1 GeV it spans about 1e5 down to
    1e-30."""
    rng = np.random.default_rng(seed)
    eg = grids["egrid"]
    rm = nodes["rmag"]
    S = rm.shape[0]
    g = rng.uniform(index[0], index[1], S)
    spec = (eg[None, :] / eg[0]) ** g[:, None]                      # [S, E]
    rad = (rm / rm[:, :1]) ** radial_power                          # [S, N]
    f = f_max * rad[:, :, None] * spec[:, None, :]
    return np.repeat(f[..., None], M, axis=3)

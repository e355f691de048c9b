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


# The benchmark workloads (rows, cols, nodes N, energies E, mu cells M, EP
# substeps per global step), shared by bench.py and the tests that check
# them at full per-stream size.
WORKLOADS = {
    # plain method, small: 24 streams x 400 x 20 x 20
    "c1": dict(rows=2, cols=2, N=400, E=20, M=20, eps=10),
    # plain method at scale: 384 streams x 2000 x 50 x 30
    "c2": dict(rows=8, cols=8, N=2000, E=50, M=30, eps=10),
    # batched form: 1536 streams whose active-node counts are drawn from a
    # seeded uniform 200..4000, as shells are injected; 64 x 32
    "c3": dict(rows=16, cols=16, N=4000, E=64, M=32, eps=10, ragged=(200, 4000)),
    # mixed-precision form on the 384 x 2000 x 50 x 30 workload with a
    # 35-decade power-law f
    "c4": dict(rows=8, cols=8, N=2000, E=50, M=30, eps=10, mixed=True),
    # learned large-step surrogate on the 384 x 2000 x 50 x 30 workload
    "c5": dict(rows=8, cols=8, N=2000, E=50, M=30, eps=10, surrogate=16),
}


def workload_params(c):
    """EpParams of the benchmark workloads: lambda = 0.1 AU (R / 1 GV)^(1/3)
    (r / 1 AU), seed spectrum E^-2 without cutoff, B_r = 2.5 nT at 1 AU,
    10 keV - 1 GeV, tDel = 1/64 day."""
    from . import config
    return config.default_params(
        rScale=0.00465047, tDel=1.0 / 64, numEpSteps=c["eps"], eMin=0.01, eMax=1000.0, lamo=0.1,
        rigidityPower=1.0 / 3.0, mfpRadialPower=1.0, mhdBAu=2.5e-5, boundaryFunctGamma=2.0,
        boundaryFunctEcutoff=1.0e30)


def workload_nodes(c, seed=20261203, cme_r=0.5):
    """Parker spiral (400 km/s, B_r 2.5 nT at 1 AU, 25-day rotation) with a
    smooth seeded 10% perturbation (correlation ~20 nodes), 2% spacing
    jitter and a Gaussian CME shell at cme_r AU (1500 km/s, density and
    field x4, 60 deg wide, 0.05 AU thick)."""
    return parker_nodes(
        c["rows"], c["cols"], c["N"], tdel_days=1.0 / 64, flow_cms=400e5, b_au=2.5e-5,
        omega=2 * math.pi * K.TAU / (25 * 86400.0), rscale=0.00465047, seed=seed, perturb=0.1,
        smooth=20, jitter=0.02,
        cme=dict(r_au=cme_r, width_au=0.05, speed_cms=1500e5, dens=4.0, ang_deg=60.0))


def workload_active_counts(c, S):
    """Per-stream active-node counts: None (all N) unless the workload is
    ragged, then a seeded uniform draw over its range."""
    if "ragged" not in c:
        return None
    lo, hi = c["ragged"]
    return np.random.default_rng(20261203).integers(lo, hi + 1, S).astype(np.int32)

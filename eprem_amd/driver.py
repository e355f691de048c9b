"""Host time loop of an analytic-wind run with every node operation on the
device.

This is the part of the reference's C host that drives the hot path
(src/eprem.c:426-744 with masCouple = 0, src/simCore.c:56-785): it builds
the cube-shell grid (initCubeCoords / initSphereCoords,
src/cubeShellInit.c:46-91, :529-822), and for every global step writes the
observer output when it is due, spawns the new inner shell from the
rotating template (rotSunAndSpawnShell, src/simCore.c:442-687), moves the
nodes with the wind, updates the MHD values on the nodes and runs the EP
update.  Spawn, ripple, node motion, updateMhd and the EP update all run
in libeprem_hip.so; the host only computes the S new shell positions per
step and downloads what an output dump writes.
"""
import math
import os

import numpy as np

from . import abi, config, constants as K, ncio, runtime

DBL_MIN = 2.2250738585072014e-308


def cube_shell_grid(rows, cols, n_nodes):
    """Template shell (index 0) and node positions of every stream,
    [S][n_nodes + 1], rScale units: the face coordinates accumulated by
    repeated +/- cell widths as initCubeCoords does, normalised onto the
    unit sphere, shells k at radius 1 + k.  Also returns the template
    colatitude and azimuth per stream."""
    S = 6 * rows * cols
    L = n_nodes + 1
    x, y, z, rm = (np.zeros((S, L)) for _ in range(4))
    zen_t, azi_t = np.zeros(S), np.zeros(S)
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
                cx, cy, cz = [(u, w, 0.5), (u2, w, -0.5), (0.5, u, w), (-0.5, u2, w), (u2, 0.5, w),
                              (u, -0.5, w)][face]
                s = (face * rows + row) * cols + col
                r = math.sqrt((cx * cx) + (cy * cy) + (cz * cz))
                cx, cy, cz = cx / r, cy / r, cz / r
                azi = math.atan2(cy, cx)
                zen = math.acos(cz / 1.0)
                zen_t[s], azi_t[s] = zen, azi
                for k in range(L):
                    rk = 1.0 + 0 + k if k > 0 else 1.0
                    rm[s, k] = rk
                    z[s, k] = rk * math.cos(zen)
                    x[s, k] = rk * math.sin(zen) * math.cos(azi)
                    y[s, k] = rk * math.sin(zen) * math.sin(azi)
    return x, y, z, rm, zen_t, azi_t


def _angles(x, y, z, rm):
    """Colatitude and azimuth in [0, 2 pi] as the output writer forms them
    (src/unifiedOutput.c:480-488), with the C library's acos/atan2."""
    zen = np.vectorize(math.acos)(z / rm)
    azi = np.vectorize(math.atan2)(y, x)
    azi = np.where(azi < 0.0, azi + 2.0 * K.PI, azi)
    azi = np.where(azi > 2.0 * K.PI, azi - 2.0 * K.PI, azi)
    return zen, azi


class ParkerRun:
    """One analytic-wind run (masCouple = 0) on the device."""

    def __init__(self, cfg):
        config.check_host_loop_options(cfg)
        self.cfg = cfg
        self.R, self.C = cfg["numRowsPerFace"], cfg["numColumnsPerFace"]
        self.N, self.E, self.M = cfg["numNodesPerStream"], cfg["numEnergySteps"], cfg["numMuSteps"]
        self.S = 6 * self.R * self.C
        self.p = config.ep_params(cfg)
        self.sol = runtime.EpSolver((self.S, self.N, self.E, self.M), self.p)

    def close(self):
        self.sol.close()

    def run(self, max_dumps=64, outdir=None):
        """Returns the dumps in code units (as the output files hold them)
        and the per-step EP diagnostics; writes obs%06d.nc files when outdir
        is given."""
        cfg, S, N, E, M = self.cfg, self.S, self.N, self.E, self.M
        p, sol = self.p, self.sol
        tdel = p.tDel
        sim_start = cfg["simStartTime"] / K.DAY
        sim_stop = cfg["simStopTime"] / K.DAY
        ep_start = cfg["epCalcStartTime"]
        out_time = cfg["simStartTime"] + 0.5 * tdel * K.DAY
        mhd = dict(mode=0, mhdUs=cfg["flowMag"] / K.C, mhdNsAu=cfg["mhdDensityAu"] / K.MHD_DENSITY_NORM,
                   mhdBsAu=cfg["mhdBAu"] / K.MHD_B_NORM, omegaSun=cfg["omegaSun"])
        init_domain = p.useBoundaryFunction > 0 and cfg["boundaryFunctionInitDomain"] > 0

        x, y, z, rm, zen_t, azi_t = cube_shell_grid(self.R, self.C, N)
        sol.set_positions({"x": x[:, 1:], "y": y[:, 1:], "z": z[:, 1:], "rmag": rm[:, 1:]})
        sol.upload_f(np.full((S, N, E, M), DBL_MIN))
        grids = sol.grids()
        zeros = np.zeros((S, N))
        fields = {k: zeros.copy() for k in ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density")}
        dbl_min_f = np.full((S, E, M), DBL_MIN)

        azi_sun = cfg["aziSunStart"]
        t_global = sim_start - N * tdel
        rciter = num_loops = 0
        sim_started = ep_init = 0
        have_mhd = False
        dumps, diags = [], []
        while True:
            rciter += 1
            if rciter == N + 1 and sim_started == 0:
                sim_started = 1
                if init_domain:
                    sol.init_seed_population()
                t_global = sim_start
            if t_global * K.DAY >= ep_start and sim_started == 1:
                sim_started = 2
            if cfg["dumpFreq"] > 0 and num_loops % cfg["dumpFreq"] == 0 and t_global * K.DAY >= out_time \
                    and len(dumps) < max_dumps:
                pos = sol.positions()
                zen, azi = _angles(pos["x"], pos["y"], pos["z"], pos["rmag"])
                d = {"time": t_global, "R": pos["rmag"], "T": zen, "P": azi,
                     "Br": fields["br"], "Bt": fields["btheta"], "Bp": fields["bphi"],
                     "Vr": fields["vr"], "Vt": fields["vtheta"], "Vp": fields["vphi"],
                     "Rho": fields["density"], "dist": sol.download_f()}
                dumps.append({k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in d.items()})
            # rotSunAndSpawnShell: the template rotates with the Sun
            azi_sun += cfg["omegaSun"] * tdel
            if azi_sun > K.PI:
                azi_sun -= 2.0 * K.PI
            sx, sy, sz = np.zeros(S), np.zeros(S), np.zeros(S)
            for s in range(S):
                a = azi_t[s] + azi_sun
                if a > K.PI:
                    a -= 2.0 * K.PI
                zn = zen_t[s]
                sx[s] = 1.0 * math.sin(zn) * math.cos(a)
                sy[s] = 1.0 * math.sin(zn) * math.sin(a)
                sz[s] = 1.0 * math.cos(zn)
            # the template's distribution: DBL_MIN until the initial seed
            # population, then the seed spectrum at its radius
            sol.spawn_shell(sx, sy, sz, np.ones(S), None if (sim_started and init_domain) else dbl_min_f)
            if have_mhd:
                sol.move_nodes(tdel)
                out = sol.update_mhd(None, **mhd)
            else:
                # before the first MHD update every node is at rest and its
                # MHD state is zero
                pos = sol.positions()
                prev = {"rold_x": pos["x"], "rold_y": pos["y"], "rold_z": pos["z"]}
                prev.update({k: zeros for k in ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density")})
                out = sol.update_mhd(prev, **mhd)
                have_mhd = True
            fields = {k: out[k] for k in fields}
            if t_global * K.DAY >= ep_start:
                if ep_init == 0:
                    ep_init = 1
                    if init_domain:
                        sol.init_seed_population()
                rc, dg = sol.step(raise_on_error=False)
                diags.append(dg)
                if cfg["seedFunctionTest"] > 0 and init_domain:
                    sol.init_seed_population()
                if rc != 0:
                    break
            t_global += tdel
            num_loops += 1
            if not t_global <= sim_stop:
                break
        res = {k: np.array([d[k] for d in dumps]) for k in
               ("time", "R", "T", "P", "Br", "Bt", "Bp", "Vr", "Vt", "Vp", "Rho", "dist")}
        res["egrid"], res["vgrid"], res["mu"] = grids["egrid"], grids["vgrid"], grids["mugrid"]
        res["diag"] = diags
        if outdir:
            os.makedirs(outdir, exist_ok=True)
            for s in range(S):
                raw = {k: res[k][:, s] for k in ("R", "T", "P", "Br", "Bt", "Bp", "Vr", "Vt", "Vp", "Rho")}
                raw["Dist"] = res["dist"][:, s][:, :, None, :, :]
                raw.update(time=res["time"], mu=res["mu"], egrid=res["egrid"], vgrid=res["vgrid"],
                           mass=np.asarray(cfg["mass"], np.float64), charge=np.asarray(cfg["charge"], np.float64))
                ncio.write_observer_file(os.path.join(outdir, ncio.stream_file_name(s)), raw, cfg["rScale"])
        return res

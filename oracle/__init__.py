"""ctypes binding of the FP64 CPU oracle (oracle/ep_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product package
(eprem_amd) never imports it.
"""
import ctypes as C
import os

import numpy as np

from eprem_amd.abi import (EP_OP_ALL, EpDiag, EpDims, EpParams, dptr, fptr, iptr)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libep_oracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle` "
                               "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        DP = C.POINTER(C.c_double)
        IP = C.POINTER(C.c_int32)
        L.or_build_grids.argtypes = [C.POINTER(EpParams), C.c_int, C.c_int] + [DP] * 8
        L.or_seed.argtypes = [C.POINTER(EpParams), C.c_double, C.c_double]
        L.or_seed.restype = C.c_double
        L.or_mfp.argtypes = [C.POINTER(EpParams), C.c_double, C.c_double, C.c_double]
        L.or_mfp.restype = C.c_double
        L.or_update_ep.argtypes = ([C.POINTER(EpDims), C.POINTER(EpParams), IP] + [DP] * 12 +
                                   [C.c_uint32, C.c_int, C.POINTER(EpDiag)])
        L.or_update_ep.restype = C.c_int
        L.or_get_stream_list.argtypes = [C.POINTER(EpParams), C.c_int, DP, DP, DP, IP, IP, DP, DP, DP]
        L.or_get_stream_list.restype = C.c_int
        L.or_focus_operator.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, DP, DP, DP]
        L.or_focus_operator.restype = None
        L.or_change_operator.argtypes = [C.POINTER(EpParams), C.c_int, C.c_int, C.c_double, DP, DP, DP]
        L.or_change_operator.restype = None
        L.or_mas_bsearch.argtypes = [C.POINTER(C.c_float), C.c_double, IP, IP, C.c_int, C.c_int]
        L.or_mas_bsearch.restype = C.c_double
        FPA = C.POINTER(C.c_float) * 7
        MPA = C.POINTER(C.c_float) * 21
        L.or_mas_interp.argtypes = ([C.c_int, C.c_double, DP, DP, DP, DP, FPA, MPA, IP, FPA, MPA, IP,
                                     C.c_double, C.c_double, C.c_double, C.c_double] + [DP] * 10 +
                                    [DP] * 7)
        L.or_observers.argtypes = ([C.POINTER(EpDims), IP, C.c_double, C.c_double, C.c_int] +
                                   [DP] * 3 + [DP] * 4 + [DP, DP, DP, DP])
        L.or_stream_flux.argtypes = [C.POINTER(EpDims), DP, DP, DP]
        L.or_move_nodes.argtypes = [C.c_int, C.c_double, C.c_double] + [DP] * 10
        L.or_move_nodes.restype = None
        FPP_ = C.POINTER(C.POINTER(C.c_float))
        L.or_move_nodes_mas.argtypes = ([C.c_int, C.c_double, C.c_double] + [DP] * 7 +
                                        [C.POINTER(DP), FPP_, FPP_, IP, FPP_, FPP_, IP] + [C.c_double] * 4)
        L.or_move_nodes_mas.restype = None
        L.or_update_mhd.argtypes = ([C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, DP, DP, DP, DP, IP,
                                     C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.POINTER(C.c_float)), IP,
                                     C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.POINTER(C.c_float)), IP,
                                     C.c_double, C.c_double, C.c_double, C.c_double] + [DP] * 10 +
                                    [DP] * 11)
        L.or_run_parker.argtypes = [C.POINTER(SimConfig), C.POINTER(SimOutput), C.c_int]
        L.or_run_parker.restype = C.c_int
        _lib = L
    return _lib


class SimConfig(C.Structure):
    _fields_ = [
        ("rows", C.c_int), ("cols", C.c_int), ("num_nodes", C.c_int),
        ("num_energies", C.c_int), ("num_mu", C.c_int),
        ("rScale", C.c_double), ("flowMag", C.c_double), ("mhdDensityAu", C.c_double),
        ("mhdBAu", C.c_double),
        ("simStartTime", C.c_double), ("simStopTime", C.c_double), ("tDel", C.c_double),
        ("numEpSteps", C.c_int),
        ("aziSunStart", C.c_double), ("omegaSun", C.c_double),
        ("dsh_min", C.c_double), ("lamo", C.c_double), ("mfpRadialPower", C.c_double),
        ("mfpBPower", C.c_double), ("rigidityPower", C.c_double),
        ("mfpType", C.c_int),
        ("eMin", C.c_double), ("eMax", C.c_double), ("mass", C.c_double), ("charge", C.c_double),
        ("dampCooling", C.c_double),
        ("boundaryFunctAmplitude", C.c_double), ("boundaryFunctXi", C.c_double),
        ("boundaryFunctGamma", C.c_double), ("boundaryFunctBeta", C.c_double),
        ("boundaryFunctEcutoff", C.c_double), ("boundaryFunctR0", C.c_double),
        ("useParallelDiffusion", C.c_int), ("useAdiabaticFocus", C.c_int),
        ("useAdiabaticChange", C.c_int), ("checkSeedPopulation", C.c_int),
        ("useEPBoundary", C.c_int), ("useBoundaryFunction", C.c_int),
        ("boundaryFunctionInitDomain", C.c_int), ("seedFunctionTest", C.c_int),
        ("dumpFreq", C.c_int),
        ("unifiedOutputTime", C.c_double), ("epCalcStartTime", C.c_double),
    ]


class SimOutput(C.Structure):
    _fields_ = ([("max_dumps", C.c_int), ("num_dumps", C.c_int), ("max_steps", C.c_int),
                 ("num_steps", C.c_int), ("time", C.POINTER(C.c_double))] +
                [(n, C.POINTER(C.c_double)) for n in
                 ("R", "T", "P", "Br", "Bt", "Bp", "Vr", "Vt", "Vp", "Rho", "dist",
                  "egrid", "vgrid", "mu")] +
                [("step_diag", C.POINTER(EpDiag)), ("raw", C.c_int)])


def build_grids(p, E, M):
    out = {k: np.zeros(E) for k in ("egrid", "vgrid", "vgrid_i", "pgrid", "rigidity")}
    out["mugrid"] = np.zeros(M)
    dl, dm = C.c_double(), C.c_double()
    lib().or_build_grids(C.byref(p), E, M, dptr(out["egrid"]), dptr(out["vgrid"]),
                         dptr(out["vgrid_i"]), dptr(out["pgrid"]), dptr(out["rigidity"]),
                         dptr(out["mugrid"]), C.byref(dl), C.byref(dm))
    out["dlnp"], out["dmu"] = dl.value, dm.value
    return out


def seed(p, energy, r):
    return lib().or_seed(C.byref(p), float(energy), float(r))


def mfp(p, rig, rmag, bmag=1.0):
    return lib().or_mfp(C.byref(p), float(rig), float(rmag), float(bmag))


def get_stream_list(p, x, y, z):
    n = len(x)
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    lst = np.zeros(n, np.int32)
    ref = np.zeros(n, np.int32)
    dsl = np.zeros(n)
    dsil = np.zeros(n)
    dmin = C.c_double()
    size = lib().or_get_stream_list(C.byref(p), n, dptr(x), dptr(y), dptr(z), iptr(lst), iptr(ref),
                                    dptr(dsl), dptr(dsil), C.byref(dmin))
    if size < 0:
        return size, None, None, None, None, dmin.value
    return size, lst[:size], ref, dsl[:size], dsil[:size], dmin.value


def update_ep(dims, p, nodes, f, n_active=None, op_flags=EP_OP_ALL, nthreads=0):
    """Run the oracle's updateEnergeticParticles on copies of the inputs.

    nodes: dict with x, y, z, rmag, bmag, dlnb, dlnn, dupar arrays [S, N].
    f: [S, N, E, M].  Returns (rc, diag dict, f_out, ds, bplus, bminus).
    """
    S, N = dims.num_streams, dims.num_nodes
    arrs = {k: np.ascontiguousarray(nodes[k], dtype=np.float64).reshape(S * N)
            for k in ("x", "y", "z", "rmag", "bmag", "dlnb", "dlnn", "dupar")}
    fo = np.ascontiguousarray(f, dtype=np.float64).copy()
    ds, bp, bm = np.zeros(S * N), np.zeros(S * N), np.zeros(S * N)
    na = None if n_active is None else np.ascontiguousarray(n_active, dtype=np.int32)
    dg = EpDiag()
    rc = lib().or_update_ep(C.byref(dims), C.byref(p), iptr(na), dptr(arrs["x"]), dptr(arrs["y"]),
                            dptr(arrs["z"]), dptr(arrs["rmag"]), dptr(arrs["bmag"]),
                            dptr(arrs["dlnb"]), dptr(arrs["dlnn"]), dptr(arrs["dupar"]), dptr(fo),
                            dptr(ds), dptr(bp), dptr(bm), op_flags, nthreads, C.byref(dg))
    return rc, dg.as_dict(), fo, ds.reshape(S, N), bp.reshape(S, N), bm.reshape(S, N)


def sim_config_from(cfg):
    """SimConfig from a resolved input-file config (eprem_amd.config.read_config)."""
    c = SimConfig()
    c.rows, c.cols = cfg["numRowsPerFace"], cfg["numColumnsPerFace"]
    c.num_nodes, c.num_energies, c.num_mu = cfg["numNodesPerStream"], cfg["numEnergySteps"], cfg["numMuSteps"]
    for k in ("rScale", "flowMag", "mhdDensityAu", "mhdBAu", "simStartTime", "simStopTime", "tDel",
              "numEpSteps", "aziSunStart", "omegaSun", "dsh_min", "lamo", "mfpRadialPower",
              "mfpBPower", "rigidityPower", "mfpType", "eMin", "eMax", "dampCooling",
              "boundaryFunctAmplitude", "boundaryFunctXi", "boundaryFunctGamma",
              "boundaryFunctBeta", "boundaryFunctEcutoff", "boundaryFunctR0",
              "useParallelDiffusion", "useAdiabaticFocus", "useAdiabaticChange",
              "checkSeedPopulation", "useEPBoundary", "useBoundaryFunction",
              "boundaryFunctionInitDomain", "seedFunctionTest", "dumpFreq"):
        setattr(c, k, cfg[k])
    c.mass, c.charge = cfg["mass"][0], cfg["charge"][0]
    c.unifiedOutputTime = -1.0
    c.epCalcStartTime = cfg["epCalcStartTime"]
    return c


def run_parker(cfg, max_dumps=64, max_steps=4096, nthreads=0, raw=False):
    """Run the analytic-wind pipeline; returns a dict of numpy outputs
    (physical units, or the code-unit values the output file holds when
    raw is set)."""
    c = sim_config_from(cfg)
    S, N, E, M = 6 * c.rows * c.cols, c.num_nodes, c.num_energies, c.num_mu
    bufs = {"time": np.zeros(max_dumps)}
    for k in ("R", "T", "P", "Br", "Bt", "Bp", "Vr", "Vt", "Vp", "Rho"):
        bufs[k] = np.zeros((max_dumps, S, N))
    bufs["dist"] = np.zeros((max_dumps, S, N, E, M))
    bufs["egrid"], bufs["vgrid"], bufs["mu"] = np.zeros(E), np.zeros(E), np.zeros(M)
    diags = (EpDiag * max_steps)()
    o = SimOutput()
    o.max_dumps, o.max_steps = max_dumps, max_steps
    o.raw = 1 if raw else 0
    for k, v in bufs.items():
        setattr(o, k, dptr(v))
    o.step_diag = C.cast(diags, C.POINTER(EpDiag))
    rc = lib().or_run_parker(C.byref(c), C.byref(o), nthreads)
    nd = o.num_dumps
    res = {k: (v[:nd] if k not in ("egrid", "vgrid", "mu") else v) for k, v in bufs.items()}
    res["rc"] = rc
    res["diag"] = [diags[i].as_dict() for i in range(o.num_steps)]
    return res


def focus_operator(alg, f, v_avg, dmu):
    """d(v f)/dmu of the focusing operator (alg 1 upwind, 3 WENO3) for
    f [E][M] and half-mesh velocities v_avg [E][M+1]."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    va = np.ascontiguousarray(v_avg, dtype=np.float64)
    E, M = f.shape
    assert va.shape == (E, M + 1)
    f1 = np.zeros((E, M))
    lib().or_focus_operator(int(alg), E, M, float(dmu), dptr(f), dptr(va), dptr(f1))
    return f1


def change_operator(p, lnf, v_avg, rmag):
    """The adiabatic-change operator of p.adiabaticChangeAlg on ln f [E][M]
    with v_avg [E+1][M] at node radius rmag."""
    lnf = np.ascontiguousarray(lnf, dtype=np.float64)
    va = np.ascontiguousarray(v_avg, dtype=np.float64)
    E, M = lnf.shape
    assert va.shape == (E + 1, M)
    f1 = np.zeros((E, M))
    lib().or_change_operator(C.byref(p), E, M, float(rmag), dptr(lnf), dptr(va), dptr(f1))
    return f1


def mas_interp(rScale, x, y, z, rmag, level0, level1, s_cor, phi_offset, mas_radial_max, mas_vmin,
               prev=None):
    """level = (fields[7] float32 arrays, meshes[21] float32 arrays)."""
    n = len(x)
    FPA = C.POINTER(C.c_float) * 7
    MPA = C.POINTER(C.c_float) * 21
    keep = []

    def lv(level):
        fields, meshes = level
        fields = [np.ascontiguousarray(a, dtype=np.float32) for a in fields]
        meshes = [np.ascontiguousarray(a, dtype=np.float32) for a in meshes]
        keep.extend(fields + meshes)
        dims = np.array([len(meshes[i]) for i in range(21)], np.int32)
        keep.append(dims)
        return FPA(*[fptr(a) for a in fields]), MPA(*[fptr(a) for a in meshes]), iptr(dims)

    f0, m0, d0 = lv(level0)
    f1, m1, d1 = lv(level1)
    xs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, rmag)]
    if prev is None:
        prev = {k: np.zeros(n) for k in ("rold_x", "rold_y", "rold_z", "br", "btheta", "bphi", "vr",
                                          "vtheta", "vphi", "density")}
    pv = [np.ascontiguousarray(prev[k], dtype=np.float64) for k in
          ("rold_x", "rold_y", "rold_z", "br", "btheta", "bphi", "vr", "vtheta", "vphi", "density")]
    outs = [np.zeros(n) for _ in range(7)]
    lib().or_mas_interp(n, rScale, *[dptr(a) for a in xs], f0, m0, d0, f1, m1, d1, s_cor,
                        phi_offset, mas_radial_max, mas_vmin, *[dptr(a) for a in pv],
                        *[dptr(a) for a in outs])
    return dict(zip(("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density"), outs))


def observers(dims, nodes, f, egrid, rscale, r_au, theta, phi, idw_p=3.0, n_active=None):
    S, N = dims.num_streams, dims.num_nodes
    E, M = dims.num_energies, dims.num_mu
    arr = {k: np.ascontiguousarray(nodes[k], dtype=np.float64).reshape(S * N) for k in ("x", "y", "z", "rmag")}
    r, t, ph = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (r_au, theta, phi))
    fo = np.ascontiguousarray(f, dtype=np.float64)
    eg = np.ascontiguousarray(egrid, dtype=np.float64)
    o = len(r)
    dist, flux = np.zeros((o, E, M)), np.zeros((o, E))
    na = None if n_active is None else np.ascontiguousarray(n_active, dtype=np.int32)
    lib().or_observers(C.byref(dims), iptr(na), rscale, idw_p, o, dptr(r), dptr(t), dptr(ph),
                       dptr(arr["x"]), dptr(arr["y"]), dptr(arr["z"]), dptr(arr["rmag"]), dptr(fo),
                       dptr(eg), dptr(dist), dptr(flux))
    return dist, flux


def stream_flux(dims, f, egrid):
    out = np.zeros((dims.num_streams, dims.num_nodes, dims.num_energies))
    lib().or_stream_flux(C.byref(dims), dptr(np.ascontiguousarray(f, dtype=np.float64)),
                         dptr(np.ascontiguousarray(egrid, dtype=np.float64)), dptr(out))
    return out


def update_mhd(nodes, prev, mode, rscale, mhdUs=0.0, mhdNsAu=0.0, mhdBsAu=0.0, omegaSun=0.0,
               level0=None, level1=None, s_cor=0.0, phi_offset=0.0, mas_radial_max=0.0, mas_vmin=0.0):
    """or_update_mhd over [S, N] node arrays; the first node of each stream
    is the stream's inner node."""
    S, N = nodes["x"].shape
    n = S * N
    xs = [np.ascontiguousarray(nodes[k], dtype=np.float64).ravel() for k in ("x", "y", "z", "rmag")]
    first = np.zeros((S, N), np.int32)
    first[:, 0] = 1
    first = first.ravel()
    names = ("rold_x", "rold_y", "rold_z", "br", "btheta", "bphi", "vr", "vtheta", "vphi", "density")
    pv = [np.ascontiguousarray(prev[k], dtype=np.float64).ravel() for k in names]
    FPP = C.POINTER(C.POINTER(C.c_float))
    keep = []

    def lv(level):
        if level is None:
            return None, None, None
        fields = [np.ascontiguousarray(a, dtype=np.float32) for a in level[0]]
        meshes = [np.ascontiguousarray(a, dtype=np.float32) for a in level[1]]
        dims = np.array([len(m) for m in meshes], np.int32)
        keep.extend(fields + meshes + [dims])
        fa = (C.POINTER(C.c_float) * 7)(*[fptr(a) for a in fields])
        ma = (C.POINTER(C.c_float) * 21)(*[fptr(a) for a in meshes])
        keep.extend([fa, ma])
        return C.cast(fa, FPP), C.cast(ma, FPP), iptr(dims)

    f0, m0, d0 = lv(level0)
    f1, m1, d1 = lv(level1)
    outs = [np.zeros(n) for _ in range(11)]
    lib().or_update_mhd(n, mode, rscale, mhdUs, mhdNsAu, mhdBsAu, omegaSun, *[dptr(a) for a in xs],
                        iptr(first), f0, m0, d0, f1, m1, d1, s_cor, phi_offset, mas_radial_max, mas_vmin,
                        *[dptr(a) for a in pv], *[dptr(a) for a in outs])
    keys = ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density", "bmag", "dupar", "dlnb", "dlnn")
    return {k: v.reshape(S, N) for k, v in zip(keys, outs)}


def move_nodes(nodes, state, tdel, rscale, first=1):
    """or_move_nodes on [S, N] arrays in place: nodes `first`.. of every
    stream move; returns the old positions (rold_x, rold_y, rold_z), which
    are the unmoved positions for nodes before `first`."""
    S, N = nodes["x"].shape
    pos = {k: np.ascontiguousarray(nodes[k], dtype=np.float64) for k in ("x", "y", "z", "rmag")}
    old = {k: pos[c].copy() for k, c in (("rold_x", "x"), ("rold_y", "y"), ("rold_z", "z"))}
    v = {k: np.ascontiguousarray(state[k], dtype=np.float64) for k in ("vr", "vtheta", "vphi")}
    n = N - first
    for s in range(S):
        args = [pos["x"][s, first:], pos["y"][s, first:], pos["z"][s, first:], pos["rmag"][s, first:],
                old["rold_x"][s, first:], old["rold_y"][s, first:], old["rold_z"][s, first:],
                v["vr"][s, first:], v["vtheta"][s, first:], v["vphi"][s, first:]]
        args = [np.ascontiguousarray(a) for a in args]
        lib().or_move_nodes(n, tdel, rscale, *[dptr(a) for a in args])
        for k, a in zip(("x", "y", "z", "rmag"), args[:4]):
            pos[k][s, first:] = a
        for k, a in zip(("rold_x", "rold_y", "rold_z"), args[4:7]):
            old[k][s, first:] = a
    nodes.update(pos)
    return old


def _mas_level(level, keep):
    """(fields[7], meshes[21]) as the C pointer tables and the dims array."""
    FPP = C.POINTER(C.POINTER(C.c_float))
    fields = [np.ascontiguousarray(a, dtype=np.float32) for a in level[0]]
    meshes = [np.ascontiguousarray(a, dtype=np.float32) for a in level[1]]
    dims = np.array([len(m) for m in meshes], np.int32)
    fa = (C.POINTER(C.c_float) * 7)(*[fptr(a) for a in fields])
    ma = (C.POINTER(C.c_float) * 21)(*[fptr(a) for a in meshes])
    keep.extend(fields + meshes + [dims, fa, ma])
    return C.cast(fa, FPP), C.cast(ma, FPP), iptr(dims)


MHD_NAMES = ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density")


def move_nodes_mas(nodes, state, tdel, rscale, level0, level1, s_cor, phi_offset, mas_radial_max, mas_vmin,
                   n_active=None, first=1):
    """or_move_nodes_mas (masMoveNodes) on [S, N] node arrays in place:
    nodes first..n_active-1 of every stream move through the MAS field.
    state: each node's previous old position (rold_x, rold_y, rold_z) and
    MHD state (MHD_NAMES), [S, N].  Returns the new old positions: the
    positions before the move (the unmoved ones for nodes before
    `first`)."""
    S, N = nodes["x"].shape
    na = np.full(S, N, np.int32) if n_active is None else np.asarray(n_active, np.int32)
    pos = {k: np.ascontiguousarray(nodes[k], dtype=np.float64) for k in ("x", "y", "z", "rmag")}
    old = {k: np.ascontiguousarray(state[k], dtype=np.float64).copy() for k in ("rold_x", "rold_y", "rold_z")}
    for k, c in (("rold_x", "x"), ("rold_y", "y"), ("rold_z", "z")):
        old[k][:, :first] = pos[c][:, :first]
    keep = []
    f0, m0, d0 = _mas_level(level0, keep)
    f1, m1, d1 = _mas_level(level1, keep)
    for s in range(S):
        n = int(na[s]) - first
        if n <= 0:
            continue
        sl = slice(first, first + n)
        args = [np.ascontiguousarray(a[s, sl]) for a in (pos["x"], pos["y"], pos["z"], pos["rmag"],
                                                         old["rold_x"], old["rold_y"], old["rold_z"])]
        pv = [np.ascontiguousarray(np.asarray(state[k], dtype=np.float64)[s, sl]) for k in MHD_NAMES]
        pa = (C.POINTER(C.c_double) * 7)(*[dptr(a) for a in pv])
        lib().or_move_nodes_mas(n, tdel, rscale, *[dptr(a) for a in args], pa, f0, m0, d0, f1, m1, d1,
                                s_cor, phi_offset, mas_radial_max, mas_vmin)
        for k, a in zip(("x", "y", "z", "rmag"), args[:4]):
            pos[k][s, sl] = a
        for k, a in zip(("rold_x", "rold_y", "rold_z"), args[4:7]):
            old[k][s, sl] = a
    nodes.update(pos)
    return old


def shell_diffuse(p, nodes, mhd, f, nbr, grids, dt, kperxkpar):
    """or_shell_diffuse on copies: nodes x,y,z,rmag [S,N], mhd br,btheta,bphi
    [S,N], f [S,N,E,M], nbr [S,4]; returns (rc, f_out, bad[4])."""
    S, N, E, M = f.shape
    xs = [np.ascontiguousarray(nodes[k], dtype=np.float64) for k in ("x", "y", "z", "rmag")]
    bs = [np.ascontiguousarray(mhd[k], dtype=np.float64) for k in ("br", "btheta", "bphi")]
    fo = np.ascontiguousarray(f, dtype=np.float64).copy()
    nb = np.ascontiguousarray(nbr, dtype=np.int32)
    rig = np.ascontiguousarray(grids["rigidity"], dtype=np.float64)
    vg = np.ascontiguousarray(grids["vgrid"], dtype=np.float64)
    bad = np.full(4, -1, np.int32)
    L = lib()
    DP, IP = C.POINTER(C.c_double), C.POINTER(C.c_int32)
    L.or_shell_diffuse.argtypes = [C.POINTER(EpParams), C.c_int, C.c_int, C.c_int, C.c_int, IP] + [DP] * 9 + \
        [C.c_double, C.c_double, DP, IP]
    rc = L.or_shell_diffuse(C.byref(p), S, N, E, M, iptr(nb), *[dptr(a) for a in xs], *[dptr(a) for a in bs],
                            dptr(rig), dptr(vg), dt, kperxkpar, dptr(fo), iptr(bad))
    return rc, fo, bad


def curl_b_over_b2(rscale, mhdBsAu, omega, x, y, z, vr):
    L = lib()
    L.or_curl_b_over_b2.argtypes = [C.c_double] * 7 + [C.POINTER(C.c_double)]
    L.or_curl_b_over_b2.restype = None
    out = (C.c_double * 3)()
    L.or_curl_b_over_b2(rscale, mhdBsAu, omega, x, y, z, vr, out)
    return np.array(out[:])


def shell_drift(p, nodes, curl, f, nbr, grids, dt):
    """or_shell_drift on copies: curl [S,N,3], f [S,N,E,M]; returns
    (rc, f_out, bad[4])."""
    S, N, E, M = f.shape
    xs = [np.ascontiguousarray(nodes[k], dtype=np.float64) for k in ("x", "y", "z")]
    cu = np.ascontiguousarray(curl, dtype=np.float64)
    fo = np.ascontiguousarray(f, dtype=np.float64).copy()
    nb = np.ascontiguousarray(nbr, dtype=np.int32)
    pg = np.ascontiguousarray(grids["pgrid"], dtype=np.float64)
    vg = np.ascontiguousarray(grids["vgrid"], dtype=np.float64)
    bad = np.full(4, -1, np.int32)
    L = lib()
    DP, IP = C.POINTER(C.c_double), C.POINTER(C.c_int32)
    L.or_shell_drift.argtypes = [C.POINTER(EpParams), C.c_int, C.c_int, C.c_int, C.c_int, IP] + [DP] * 6 + \
        [C.c_double, DP, IP]
    rc = L.or_shell_drift(C.byref(p), S, N, E, M, iptr(nb), *[dptr(a) for a in xs], dptr(cu), dptr(pg), dptr(vg),
                          dt, dptr(fo), iptr(bad))
    return rc, fo, bad

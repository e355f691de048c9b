"""GPU parity: the HIP library (through the C ABI) against the FP64 oracle.

Bar: 1e-10 relative on every cell of f (the arithmetic keeps the
reference's operation order without FMA contraction, so the observed
error is the device libm's exp/log/pow rounding, ~1e-15), identical
integer diagnostics (subcycle counts) and identical error codes/locations.
"""
import numpy as np
import pytest

import oracle
from eprem_amd import abi, config, runtime, workloads

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def params(**kw):
    base = dict(numEpSteps=4, tDel=1.0 / 64, eMin=0.01, eMax=1000.0, lamo=0.1, mfpRadialPower=1.0,
                boundaryFunctEcutoff=1.0e30)
    base.update(kw)
    return config.default_params(**base)


def run_both(p, nodes, E, M, f0=None, n_active=None, op_flags=abi.EP_OP_ALL, steps=1, tiling=None,
             noise=0.3, seed=2, precision=0):
    dims = workloads.dims_of(nodes, E, M)
    sol = runtime.EpSolver(dims, p)
    try:
        if precision:
            sol.set_precision(precision)
        if tiling:
            sol.set_tuning(*tiling)
        grids = sol.grids()
        if f0 is None:
            f0 = workloads.seeded_f(p, grids, nodes, noise=noise, seed=seed)
        sol.set_nodes(nodes, n_active)
        sol.upload_f(f0)
        gd = []
        for _ in range(steps):
            rc, dg = sol.step(op_flags, raise_on_error=False)
            gd.append((rc, dg))
            if rc:
                break
        fg = sol.download_f()
        outs = sol.node_outputs()
    finally:
        sol.close()
    fo = f0.copy()
    od = []
    for _ in range(steps):
        rc, do, fo, ds, bp, bm = oracle.update_ep(dims, p, nodes, fo, n_active=n_active, op_flags=op_flags)
        od.append((rc, do))
        if rc:
            break
    return fg, fo, gd, od, outs, (ds, bp, bm)


def max_rel(a, b, n_active=None):
    if n_active is not None:
        mask = np.zeros(a.shape[:2], bool)
        for s, n in enumerate(n_active):
            mask[s, :n] = True
        a, b = a[mask], b[mask]
    return float(np.max(np.abs(a - b) / np.abs(b)))


def active_mask(shape, n_active):
    mask = np.zeros(shape, bool)
    for s, n in enumerate(n_active):
        mask[s, :n] = True
    return mask


def ulp_spread(dims, p, nodes, f0, n_active, steps, samples=4, seed=11):
    """Per-cell largest move of the oracle's answer over runs whose node
    inputs (coordinates, |B|, the MHD differences) each differ from the
    given ones by about one ulp."""
    rng = np.random.default_rng(seed)
    fo = f0
    for _ in range(steps):
        fo = oracle.update_ep(dims, p, nodes, fo, n_active=n_active)[2]
    spread = np.zeros_like(fo)
    for _ in range(samples):
        nd = {k: (v * (1.0 + 2.2e-16 * rng.standard_normal(v.shape))
                  if isinstance(v, np.ndarray) and v.dtype == np.float64 else v)
              for k, v in nodes.items()}
        fp = f0
        for _ in range(steps):
            fp = oracle.update_ep(dims, p, nd, fp, n_active=n_active)[2]
        spread = np.maximum(spread, np.abs(fp - fo))
    return spread


def assert_diag_equal(gd, od):
    assert len(gd) == len(od)
    for (rg, dg), (ro, do) in zip(gd, od):
        assert rg == ro, (abi.ERR_NAMES.get(rg), abi.ERR_NAMES.get(ro))
        for k in ("max_subcycles_change", "max_subcycles_focusing", "max_nsteps_stream"):
            assert dg[k] == do[k], (k, dg[k], do[k])
        assert dg["min_tau"] == pytest.approx(do["min_tau"], rel=1e-14)
        assert dg["cell_updates"] == do["cell_updates"]


def test_dpp_lane_layout():
    x, shr, shl = runtime.selftest_dpp()
    np.testing.assert_array_equal(shr[1:], x[:-1])
    assert shr[0] == -1.0
    np.testing.assert_array_equal(shl[:-1], x[1:])
    assert shl[63] == -2.0


@pytest.mark.parametrize("E,M", [(12, 6), (20, 20), (8, 5)])
def test_grids_bit_exact(E, M):
    p = params()
    sol = runtime.EpSolver((6, 8, E, M), p)
    try:
        g = sol.grids()
    finally:
        sol.close()
    o = oracle.build_grids(p, E, M)
    for k in ("egrid", "vgrid", "mugrid", "rigidity", "dlnp", "dmu"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)


@pytest.mark.parametrize("case", [
    dict(rows=1, cols=1, N=64, E=12, M=6),
    dict(rows=1, cols=1, N=80, E=20, M=20),
    dict(rows=1, cols=1, N=70, E=9, M=7),
    dict(rows=1, cols=1, N=50, E=50, M=30),
    dict(rows=1, cols=1, N=40, E=64, M=32),
    dict(rows=2, cols=2, N=60, E=12, M=6),
])
def test_step_parity(case):
    p = params()
    nodes = workloads.parker_nodes(case["rows"], case["cols"], case["N"], seed=1, perturb=0.2, jitter=0.2)
    fg, fo, gd, od, outs, oouts = run_both(p, nodes, case["E"], case["M"])
    assert_diag_equal(gd, od)
    assert max_rel(fg, fo) <= RTOL
    for a, b in zip(outs, oouts):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("flags", [abi.EP_OP_STREAM, abi.EP_OP_FOCUS, abi.EP_OP_CHANGE,
                                   abi.EP_OP_SEEDCHECK, abi.EP_OP_STREAM | abi.EP_OP_SEEDCHECK])
def test_single_operator_parity(flags):
    p = params()
    nodes = workloads.parker_nodes(1, 1, 64, seed=3, perturb=0.3, jitter=0.1)
    fg, fo, gd, od, _, _ = run_both(p, nodes, 12, 8, op_flags=flags)
    assert_diag_equal(gd, od)
    assert max_rel(fg, fo) <= RTOL


@pytest.mark.parametrize("tiling", [(64, 1), (64, 20), (128, 8), (256, 32), (512, 64), (512, 200),
                                    (128, 1, 2), (128, 20, 2), (256, 7, 2), (512, 32, 2), (1024, 64, 2),
                                    (1024, 400, 2), (256, 32, 1, 0), (512, 32, 2, 0), (1024, 64, 2, 0)])
def test_streaming_tiling_invariance(tiling):
    """Results do not depend on the chunk size or subcycles per pass."""
    p = params(numEpSteps=2)
    nodes = workloads.parker_nodes(1, 1, 300, seed=4, perturb=0.2, jitter=0.3)
    fg, fo, gd, od, _, _ = run_both(p, nodes, 10, 6, tiling=tiling)
    assert_diag_equal(gd, od)
    assert max_rel(fg, fo) <= RTOL


def test_thinning_parity():
    """Nodes closer than dsh_min are thinned (GetStreamList) and copied
    from their reference node after streaming."""
    p = params()
    nodes = workloads.parker_nodes(1, 1, 120, seed=5, perturb=0.1, thin_frac=0.15)
    fg, fo, gd, od, _, _ = run_both(p, nodes, 10, 6)
    assert_diag_equal(gd, od)
    assert max_rel(fg, fo) <= RTOL


def test_ragged_streams_parity():
    """Per-stream active node counts (batched form)."""
    p = params()
    nodes = workloads.parker_nodes(1, 1, 200, seed=6, perturb=0.2, jitter=0.1)
    n_active = np.array([200, 37, 150, 3, 99, 128], np.int32)
    fg, fo, gd, od, _, _ = run_both(p, nodes, 10, 6, n_active=n_active)
    assert_diag_equal(gd, od)
    assert max_rel(fg, fo, n_active) <= RTOL


def test_multi_step_parity():
    p = params(numEpSteps=3)
    nodes = workloads.parker_nodes(1, 1, 64, seed=7, perturb=0.2, jitter=0.2)
    fg, fo, gd, od, _, _ = run_both(p, nodes, 12, 8, steps=3)
    assert_diag_equal(gd, od)
    assert max_rel(fg, fo) <= RTOL


@pytest.mark.parametrize("kw", [dict(useEPBoundary=0), dict(mfpType=2), dict(dampCooling=0.5),
                                dict(checkSeedPopulation=0), dict(lamo=1.0e-4)])
def test_option_parity(kw):
    p = params(**kw)
    nodes = workloads.parker_nodes(1, 1, 64, seed=8, perturb=0.3, jitter=0.2)
    fg, fo, gd, od, _, _ = run_both(p, nodes, 12, 8)
    assert_diag_equal(gd, od)
    assert max_rel(fg, fo) <= RTOL


def test_nan_error_location():
    """A NaN in f panics in the reference (checkNaN); the library returns
    EP_ERR_NAN at the same (stream, node, energy, mu)."""
    p = params()
    nodes = workloads.parker_nodes(1, 1, 64, seed=9, perturb=0.2)
    dims = workloads.dims_of(nodes, 12, 8)
    g = oracle.build_grids(p, 12, 8)
    f0 = workloads.seeded_f(p, g, nodes, noise=0.2, seed=1)
    f0[3, 17, 5, 2] = np.nan
    fg, fo, gd, od, _, _ = run_both(p, nodes, 12, 8, f0=f0)
    (rg, dg), (ro, do) = gd[0], od[0]
    assert rg == ro == abi.EP_ERR_NAN
    for k in ("err_stream", "err_node", "err_energy", "err_mu"):
        assert dg[k] == do[k], (k, dg[k], do[k])


def test_ds_underflow_error():
    """Two coincident kept nodes give ds = 0: FindSegmentLengths panics.
    With any positive dsh_min GetStreamList thins one of two coincident
    nodes, so the test disables thinning with a negative dsh_min (outside
    the input file's accepted range, but the same for both sides)."""
    p = params(dsh_min=-1.0)
    nodes = workloads.parker_nodes(1, 1, 32, seed=10)
    for k in ("x", "y", "z", "rmag"):
        nodes[k][2, 11] = nodes[k][2, 10]
    fg, fo, gd, od, _, _ = run_both(p, nodes, 8, 6)
    assert gd[0][0] == od[0][0] == abi.EP_ERR_DS_UNDERFLOW
    assert gd[0][1]["err_stream"] == od[0][1]["err_stream"] == 2


@pytest.mark.parametrize("sched,graph", [(0, True), (1, True), (1, False), (2, False)])
def test_preparation_error_leaves_f_unchanged(sched, graph):
    """The reference panics in FindSegmentLengths before any substep.  With
    the persistent schedule the host does not look at the preparation before
    launching the substeps: the error is found on the device, every substep
    kernel skips its work, and the step reports the error afterwards with f
    as it was."""
    p = params(dsh_min=-1.0)
    nodes = workloads.parker_nodes(1, 1, 32, seed=10)
    for k in ("x", "y", "z", "rmag"):
        nodes[k][2, 11] = nodes[k][2, 10]
    dims = workloads.dims_of(nodes, 8, 6)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_tuning(384, 16, 3, 1)
        sol.set_schedule(sched)
        sol.set_graph(graph)
        f0 = workloads.seeded_f(p, sol.grids(), nodes, noise=0.3, seed=3)
        sol.set_nodes(nodes)
        sol.upload_f(f0)
        rc, dg = sol.step(raise_on_error=False)
        fg = sol.download_f()
        # a good step afterwards runs normally
        good = workloads.parker_nodes(1, 1, 32, seed=10)
        sol.set_nodes(good)
        rc2, _ = sol.step(raise_on_error=False)
    finally:
        sol.close()
    assert rc == abi.EP_ERR_DS_UNDERFLOW and dg["err_stream"] == 2
    np.testing.assert_array_equal(fg, f0)
    assert rc2 == 0


@pytest.mark.parametrize("levels_differ", [False, True])
def test_mas_interp_parity(levels_differ):
    """masTriLinear/masWind on the device vs the oracle: nodes inside and
    outside masRadialMax, two time levels blended by s_cor, phi offset.
    Some fields sit on staggered meshes (as in MAS) and, in the second
    case, the second time level has meshes of its own for one field (the
    device reuses a search only while the mesh is the same)."""
    from eprem_amd import mas_synth
    rng = np.random.default_rng(11)
    r, t, ph = mas_synth.meshes(nr=40, nt=30, nphi=50, rmin=1.0, rmax=30.0)
    lv0 = mas_synth.level(r, t, ph, cme_r=8.0, seed=3)
    lv1 = mas_synth.level(r, t, ph, cme_r=12.0, seed=3)
    # staggered meshes for some fields, as in MAS
    m0 = list(lv0[1])
    m0[0] = (r * 1.01).astype(np.float32)
    m0[4] = (t + 0.01).astype(np.float32)
    m1 = list(m0)
    if levels_differ:
        m1[9] = (r * 0.995).astype(np.float32)
        m1[11] = (ph + 0.02).astype(np.float32)
        m1[13] = m0[4].copy()  # same points as another slot of the axis
    lv0 = (lv0[0], m0)
    lv1 = (lv1[0], m1)
    nodes = workloads.parker_nodes(1, 1, 60, tdel_days=1.0 / 512, rscale=0.00465047, seed=12, jitter=0.3)
    p = params(rScale=0.00465047)
    dims = workloads.dims_of(nodes, 4, 4)
    prev = mas_synth.prev_from_nodes(nodes, p.rScale)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_nodes(nodes)
        sol.mas_set_level(0, *lv0)
        sol.mas_set_level(1, *lv1)
        out = sol.interp_mhd(0.37, 0.8, 4.0, 50.0e5, prev)
    finally:
        sol.close()
    S, N = nodes["x"].shape
    ref = oracle.mas_interp(p.rScale, nodes["x"].ravel(), nodes["y"].ravel(), nodes["z"].ravel(),
                            nodes["rmag"].ravel(), lv0, lv1, 0.37, 0.8, 4.0, 50.0e5,
                            {k: v.ravel() for k, v in prev.items()})
    inside = (nodes["rmag"].ravel() * p.rScale) <= 4.0 * (6.96e10 / 1.495978707e13)
    assert 0 < inside.sum() < inside.size
    for k in ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density"):
        np.testing.assert_allclose(out[k].ravel(), ref[k], rtol=1e-13, atol=0, err_msg=k)


def test_mas_interp_real_mas_wind_data():
    """masTriLinear on the device vs the oracle with the reference's own
    MAS input (testsuite/mas_wind levels 1 and 2, seven fields on their own
    staggered meshes, tests/golden/mas_grids): nodes on random radial lines
    from 1.01 to 240 Rs (past the outer mesh point, where the radial weight
    clamps), with a phi offset that wraps, V_r clamped at masVmin, and the
    nodes beyond masRadialMax taking masWind."""
    import os

    from eprem_amd import constants as K
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mas_grids",
                             "mas_wind_levels.npz"))
    names = ("bp", "bt", "br", "vp", "vt", "vr", "rho")
    lv = [([z["%d_%s_data" % (l, f)] for f in names], [z["%d_%s_%s" % (l, f, a)] for f in names for a in "rtp"])
          for l in (1, 2)]
    rng = np.random.default_rng(31)
    S, N = 6, 200
    rscale = 0.00465047
    th = rng.uniform(0.05, np.pi - 0.05, S)
    ph = rng.uniform(-np.pi, np.pi, S)
    r_rs = np.exp(np.linspace(np.log(1.01), np.log(240.0), N))[None, :] * rng.uniform(0.98, 1.02, (S, N))
    rau = r_rs * K.RSAU
    nodes = {"x": rau * (np.sin(th) * np.cos(ph))[:, None] / rscale,
             "y": rau * (np.sin(th) * np.sin(ph))[:, None] / rscale,
             "z": rau * np.cos(th)[:, None] / rscale}
    nodes["rmag"] = np.sqrt(nodes["x"] ** 2 + nodes["y"] ** 2 + nodes["z"] ** 2)
    for k in ("bmag", "dlnb", "dlnn", "dupar"):
        nodes[k] = np.ones((S, N))
    from eprem_amd import mas_synth
    prev = mas_synth.prev_from_nodes(nodes, rscale)
    p = params(rScale=rscale)
    dims = workloads.dims_of(nodes, 4, 4)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_nodes(nodes)
        sol.mas_set_level(0, *lv[0])
        sol.mas_set_level(1, *lv[1])
        out = sol.interp_mhd(0.61, 2.9, 200.0, 150.0e5, prev)
    finally:
        sol.close()
    ref = oracle.mas_interp(rscale, nodes["x"].ravel(), nodes["y"].ravel(), nodes["z"].ravel(),
                            nodes["rmag"].ravel(), lv[0], lv[1], 0.61, 2.9, 200.0, 150.0e5,
                            {k: v.ravel() for k, v in prev.items()})
    inside = r_rs.ravel() <= 200.0
    assert 0 < inside.sum() < inside.size
    assert np.any(ref["vr"] == 150.0e5 / K.C)  # the clamp is exercised
    # the device's acos/atan2 of a node position can differ from glibc's by
    # an ulp, which moves the trilinear weights by ~1e-16: on a value that
    # nearly cancels between its corners (transverse components of the MAS
    # wind) that is a large relative change, so the bound is 1e-13 of each
    # component's scale
    for k in ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density"):
        np.testing.assert_allclose(out[k].ravel(), ref[k], rtol=1e-13, atol=1e-13 * np.abs(ref[k]).max(),
                                   err_msg=k)


def test_observers_and_flux_parity():
    """Point-observer projection, inverse-distance blend and flux on the
    device equal the oracle's (same summation order over streams).  The
    weights come from the device's sin/cos/pow, which may differ from glibc
    by an ulp, so the blend is held to 1e-13 relative; the stream flux has
    no transcendental and is bit-equal."""
    p = params()
    nodes = workloads.parker_nodes(1, 1, 300, seed=13, perturb=0.1, smooth=10)
    dims = workloads.dims_of(nodes, 10, 8)
    g = oracle.build_grids(p, 10, 8)
    f0 = workloads.seeded_f(p, g, nodes, noise=0.3, seed=4)
    r_au = np.array([0.3, 1.0, 0.999, 5.0])
    th = np.array([1.571, 0.4, 2.9, 1.0])
    ph = np.array([3.142, -1.0, 6.5, 0.2])
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_nodes(nodes)
        sol.upload_f(f0)
        dist, flux = sol.observe(r_au, th, ph, idw_p=3.0)
        sflux = sol.stream_flux()
    finally:
        sol.close()
    d_o, f_o = oracle.observers(dims, nodes, f0, g["egrid"], p.rScale, r_au, th, ph, 3.0)
    np.testing.assert_allclose(dist, d_o, rtol=1e-13, atol=0)
    np.testing.assert_allclose(flux, f_o, rtol=1e-13, atol=0)
    np.testing.assert_array_equal(sflux, oracle.stream_flux(dims, f0, g["egrid"]))


@pytest.mark.parametrize("mode", [0, 1])
def test_update_mhd_parity(mode):
    """updateMhd on the device (analytic wind or MAS) equals the oracle,
    and its |B| and derivative terms feed ep_step without a host copy."""
    from eprem_amd import constants as K, mas_synth
    p = params(rScale=0.00465047)
    nodes = workloads.parker_nodes(1, 1, 80, tdel_days=1.0 / 256, rscale=0.00465047, seed=14, jitter=0.2)
    prev = mas_synth.prev_from_nodes(nodes, p.rScale)
    prev = {k: v * (1.0 + 0.05 * np.sin(np.arange(v.size).reshape(v.shape))) for k, v in prev.items()}
    kw = dict(mhdUs=400e5 / K.C, mhdNsAu=8.3, mhdBsAu=2.5e-5 / K.MHD_B_NORM, omegaSun=0.0014)
    lv0 = lv1 = None
    if mode == 1:
        r, t, ph = mas_synth.meshes(nr=40, nt=30, nphi=50)
        lv0 = mas_synth.level(r, t, ph, cme_r=6.0, seed=5)
        lv1 = mas_synth.level(r, t, ph, cme_r=9.0, seed=5)
        kw.update(s_cor=0.42, phi_offset=0.3, mas_radial_max=10.0, mas_vmin=50e5)
    dims = workloads.dims_of(nodes, 8, 6)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_positions(nodes)
        if mode == 1:
            sol.mas_set_level(0, *lv0)
            sol.mas_set_level(1, *lv1)
        out = sol.update_mhd(prev, mode=mode, **kw)
        g = sol.grids()
        f0 = workloads.seeded_f(p, g, nodes, noise=0.2, seed=3)
        sol.upload_f(f0)
        _, dg = sol.step()
        fg = sol.download_f()
    finally:
        sol.close()
    ref = oracle.update_mhd(nodes, prev, mode, p.rScale, level0=lv0, level1=lv1, **kw)
    for k in ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density"):
        np.testing.assert_allclose(out[k], ref[k], rtol=1e-13, atol=0, err_msg=k)
    full = dict(nodes)
    for k in ("bmag", "dlnb", "dlnn", "dupar"):
        full[k] = ref[k]
    rc, do, fo, _, _, _ = oracle.update_ep(dims, p, full, f0)
    assert rc == 0
    assert dg["max_subcycles_focusing"] == do["max_subcycles_focusing"]
    assert max_rel(fg, fo) <= RTOL


def test_graph_replays_match_individual_launches():
    """The captured substep graph is replayed on every later step; each
    replay, with eager library work and host synchronisation between the
    replays, must give exactly what individual launches give."""
    p = params(numEpSteps=3)
    nodes = workloads.parker_nodes(1, 1, 90, seed=21, perturb=0.2, jitter=0.2)
    dims = workloads.dims_of(nodes, 12, 8)
    outs = {}
    for graph in (False, True):
        sol = runtime.EpSolver(dims, p)
        try:
            sol.set_graph(graph)
            f0 = workloads.seeded_f(p, sol.grids(), nodes, noise=0.3, seed=5)
            sol.set_nodes(nodes)
            sol.upload_f(f0)
            seq = []
            for _ in range(4):
                _, dg = sol.step()
                seq.append(sol.download_f())
                sol.stream_flux()  # eager kernels between replays
                if graph:
                    assert dg["stream_ms"] == 0.0 and dg["device_ms"] > 0.0
                else:
                    assert dg["stream_ms"] > 0.0
            outs[graph] = seq
        finally:
            sol.close()
    for a, b in zip(outs[False], outs[True]):
        np.testing.assert_array_equal(a, b)
    # and the first step agrees with the oracle
    fo = workloads.seeded_f(p, oracle.build_grids(p, 12, 8), nodes, noise=0.3, seed=5)
    _, _, fo, _, _, _ = oracle.update_ep(dims, p, nodes, fo)
    assert max_rel(outs[True][0], fo) <= RTOL


@pytest.mark.parametrize("case", [
    dict(N=300, E=10, M=30, thin=0.0, ragged=False, precision=0, tiling=(512, 16, 2)),
    dict(N=300, E=10, M=30, thin=0.0, ragged=False, precision=2, tiling=(256, 9, 2)),
    dict(N=120, E=10, M=6, thin=0.15, ragged=False, precision=0, tiling=(128, 5, 2)),
    dict(N=200, E=10, M=7, thin=0.0, ragged=True, precision=0, tiling=(256, 11, 2)),
    dict(N=200, E=8, M=16, thin=0.1, ragged=True, precision=2, tiling=(384, 30, 2)),
    dict(N=257, E=8, M=32, thin=0.0, ragged=False, precision=0, tiling=(1024, 100, 2)),
])
def test_two_nodes_per_lane_parity(case):
    """The streaming kernel with two kept nodes per lane (coefficients held
    in registers for M <= 16, formed per subcycle above) against the
    oracle: thinned and ragged streams, odd and even M, both FP64 modes,
    several steps."""
    p = params(numEpSteps=2)
    nodes = workloads.parker_nodes(1, 1, case["N"], seed=61, perturb=0.2, jitter=0.2, thin_frac=case["thin"])
    n_active = None
    if case["ragged"]:
        n_active = np.array([case["N"], 37, 150, 3, 99, 128], np.int32)
    fg, fo, gd, od, _, _ = run_both(p, nodes, case["E"], case["M"], n_active=n_active, steps=2,
                                    tiling=case["tiling"], precision=case["precision"])
    assert_diag_equal(gd, od)
    assert max_rel(fg, fo, n_active) <= RTOL


@pytest.mark.parametrize("case", [
    dict(N=300, E=10, M=30, thin=0.0, ragged=False, precision=0, tiling=(768, 48, 3)),
    dict(N=300, E=10, M=30, thin=0.0, ragged=False, precision=2, tiling=(384, 9, 3)),
    dict(N=120, E=10, M=6, thin=0.15, ragged=False, precision=0, tiling=(192, 5, 3)),
    dict(N=200, E=10, M=7, thin=0.0, ragged=True, precision=0, tiling=(576, 11, 3)),
    dict(N=200, E=8, M=16, thin=0.1, ragged=True, precision=2, tiling=(1152, 30, 3)),
    dict(N=257, E=8, M=32, thin=0.0, ragged=False, precision=0, tiling=(1536, 100, 3)),
    dict(N=150, E=12, M=30, thin=0.0, ragged=False, precision=2, tiling=(768, 24, 3, 1)),
    # several chunks per slab (interior W - 2h below the node count): the
    # halos cross chunk boundaries and only interiors are stored
    dict(N=300, E=10, M=30, thin=0.0, ragged=False, precision=0, tiling=(192, 8, 3, 0)),
    dict(N=300, E=10, M=30, thin=0.0, ragged=False, precision=2, tiling=(192, 8, 3, 1)),
    dict(N=300, E=10, M=30, thin=0.15, ragged=False, precision=2, tiling=(192, 12, 3, 1)),
    dict(N=300, E=10, M=17, thin=0.1, ragged=True, precision=0, tiling=(192, 8, 3, 0)),
    dict(N=700, E=8, M=30, thin=0.0, ragged=True, precision=2, tiling=(384, 40, 3, 1)),
])
def test_three_nodes_per_lane_parity(case):
    """The streaming kernel with three kept nodes per lane against the
    oracle: thinned and ragged streams, odd and even M, both FP64 modes,
    chunks of one to eight waves, one to several chunks per slab, focusing
    in the last pass, two steps."""
    p = params(numEpSteps=2)
    nodes = workloads.parker_nodes(1, 1, case["N"], seed=62, perturb=0.2, jitter=0.2, thin_frac=case["thin"])
    n_active = None
    if case["ragged"]:
        n_active = np.array([case["N"], 37, 150, 3, 99, 128], np.int32)
    fg, fo, gd, od, _, _ = run_both(p, nodes, case["E"], case["M"], n_active=n_active, steps=2,
                                    tiling=case["tiling"], precision=case["precision"])
    assert_diag_equal(gd, od)
    assert max_rel(fg, fo, n_active) <= RTOL


@pytest.mark.parametrize("algs", [(2, 1), (3, 1), (1, 2), (1, 3), (2, 2), (3, 3)])
@pytest.mark.parametrize("case", [
    dict(N=80, E=12, M=8, thin=0.0, precision=0),
    dict(N=60, E=20, M=30, thin=0.0, precision=2),
    dict(N=70, E=9, M=7, thin=0.15, precision=0),
])
def test_rk3_operator_parity(algs, case):
    """AdiabaticFocusing / AdiabaticChange with the three-stage Runge-Kutta
    integrator and the upwind (2) or WENO3 (3) operator against the oracle,
    with streaming on, odd and even mu counts, thinned streams, two steps."""
    falg, calg = algs
    # The WENO3 focusing operator subtracts alpha from the split fluxes
    # (LP = f v / 2 - alpha, src/energeticParticles.c:2021-2025) instead of
    # multiplying it into them.  Two consequences for a parity test:
    # - it carries no dissipation, and over hundreds of subcycles it
    #   amplifies rounding without bound (at tDel = 1/64 the oracle's own
    #   answer moves by 1e292 under one ulp of change in its inputs), so
    #   its cases run at a step of a few subcycles;
    # - the alphas, which scale with the velocities and not with f, cancel
    #   in every face flux, and the operator drives some cells to zero by
    #   cancellation (then to the floor); pitch-angle scattering spreads
    #   what is left of them along the row.  There the answer depends on
    #   the last ulps of the node inputs: one ulp of change in the node
    #   coordinates moves the oracle's own answer by up to 2e-5 in such
    #   cells.  Device logarithms and fused multiply-adds differ from the
    #   host's by about that much, so each cell is held to RTOL plus four
    #   times the oracle's own spread over runs whose node inputs differ
    #   by an ulp, and the share of cells where that spread exceeds RTOL
    #   must stay small.
    weno_focus = falg == 3
    tdel, perturb = (1.0 / 25600, 0.05) if weno_focus else (1.0 / 64, 0.2)
    p = params(numEpSteps=2, adiabaticFocusAlg=falg, adiabaticChangeAlg=calg, tDel=tdel)
    nodes = workloads.parker_nodes(1, 1, case["N"], seed=65, perturb=perturb, jitter=0.2, thin_frac=case["thin"])
    n_active = np.array([case["N"], 37, case["N"] - 5, 12, 40, 3], np.int32)
    fg, fo, gd, od, _, _ = run_both(p, nodes, case["E"], case["M"], n_active=n_active, steps=2,
                                    precision=case["precision"])
    assert_diag_equal(gd, od)
    if od[-1][0] != 0:
        # both report the same failure (WENO3 change of WENO3-focused rows
        # overflows its smoothness weights to NaN); f is not defined after it
        return
    if not weno_focus:
        assert max_rel(fg, fo, n_active) <= RTOL
        return
    dims = workloads.dims_of(nodes, case["E"], case["M"])
    f0 = workloads.seeded_f(p, oracle.build_grids(p, case["E"], case["M"]), nodes, noise=0.3, seed=2)
    spread = ulp_spread(dims, p, nodes, f0, n_active, steps=2)
    mask = active_mask(fo.shape, n_active)
    # the envelope stays narrow: the bound below is RTOL in most cells and
    # the median cell's envelope is at the 1e-13 level
    rel_spread = (spread / np.abs(fo))[mask]
    assert np.mean(rel_spread > RTOL) < 0.3
    assert np.median(rel_spread) < 1e-12
    excess = np.abs(fg - fo) / (RTOL * np.abs(fo) + 4.0 * spread)
    assert float(np.max(excess[mask])) <= 1.0


def test_rk3_operator_limits():
    """WENO3 needs four cells per line and the RK3 kernels FP64 storage:
    both are refused on the host before any device work."""
    nodes = workloads.parker_nodes(1, 1, 20, seed=66)
    with pytest.raises(runtime.EpError) as ei:
        runtime.EpSolver(workloads.dims_of(nodes, 8, 3), params(adiabaticFocusAlg=3))
    assert ei.value.code == abi.EP_ERR_UNSUPPORTED
    with pytest.raises(runtime.EpError) as ei:
        runtime.EpSolver(workloads.dims_of(nodes, 3, 8), params(adiabaticChangeAlg=3))
    assert ei.value.code == abi.EP_ERR_UNSUPPORTED
    sol = runtime.EpSolver(workloads.dims_of(nodes, 8, 6), params(adiabaticChangeAlg=2))
    try:
        with pytest.raises(runtime.EpError) as ei:
            sol.set_precision(1)
        assert ei.value.code == abi.EP_ERR_UNSUPPORTED
    finally:
        sol.close()


@pytest.mark.parametrize("sched", [0, 1, 2])
@pytest.mark.parametrize("case", [
    # two chunks per slab, several passes
    dict(N=300, E=10, M=30, thin=0.0, ragged=False, precision=0, tiling=(192, 8, 3, 0)),
    dict(N=300, E=10, M=17, thin=0.1, ragged=True, precision=2, tiling=(256, 12, 2, 1)),
    dict(N=700, E=8, M=30, thin=0.0, ragged=True, precision=2, tiling=(384, 40, 3, 1)),
    dict(N=300, E=10, M=20, thin=0.1, ragged=True, precision=2, tiling=(256, 8, 1, 1)),
    dict(N=300, E=10, M=30, thin=0.0, ragged=False, precision=0, tiling=(192, 12, 1, 0)),
    # every stream fits in one chunk: all subcycles in one pass, no halo
    dict(N=250, E=12, M=20, thin=0.0, ragged=True, precision=2, tiling=(512, 16, 1, 1)),
    dict(N=120, E=12, M=20, thin=0.0, ragged=False, precision=0, tiling=(384, 16, 3, 1)),
    dict(N=250, E=9, M=7, thin=0.15, ragged=True, precision=0, tiling=(256, 4, 2, 0)),
])
def test_streaming_schedules_parity(sched, case):
    """One launch per pass (0), one persistent launch per substep whose
    blocks take whole slabs from a queue (1) and its unbatched form, one
    launch per stream (2), against the oracle: multi-chunk and one-chunk
    slabs, thinned and ragged streams, both FP64 modes, two steps."""
    p = params(numEpSteps=2)
    nodes = workloads.parker_nodes(1, 1, case["N"], seed=63, perturb=0.2, jitter=0.2, thin_frac=case["thin"])
    n_active = None
    if case["ragged"]:
        n_active = np.array([case["N"], 37, 150, 3, 99, 128], np.int32)
    dims = workloads.dims_of(nodes, case["E"], case["M"])
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_precision(case["precision"])
        sol.set_tuning(*case["tiling"])
        sol.set_schedule(sched)
        assert sol.schedule() == sched
        f0 = workloads.seeded_f(p, sol.grids(), nodes, noise=0.3, seed=8)
        sol.set_nodes(nodes, n_active)
        sol.upload_f(f0)
        gd = [sol.step(raise_on_error=False) for _ in range(2)]
        fg = sol.download_f()
    finally:
        sol.close()
    fo, od = f0, []
    for _ in range(2):
        rc, do, fo, _, _, _ = oracle.update_ep(dims, p, nodes, fo, n_active=n_active)
        od.append((rc, do))
    assert_diag_equal(gd, od)
    assert max_rel(fg, fo, n_active) <= RTOL
    if sched:
        assert gd[0][1]["stream_launches"] == (2 if sched == 1 else 2 * dims.num_streams)


def mixed_precision_schedules(tiling):
    """Two steps in precision mode 1 under each streaming schedule; f of
    the active nodes per schedule."""
    p = params(numEpSteps=2)
    nodes = workloads.parker_nodes(1, 1, 300, seed=65, perturb=0.2, jitter=0.2, thin_frac=0.1)
    n_active = np.array([300, 37, 150, 3, 99, 128], np.int32)
    dims = workloads.dims_of(nodes, 10, 20)
    out = []
    for sched in (0, 1, 2):
        sol = runtime.EpSolver(dims, p)
        try:
            sol.set_precision(1)
            sol.set_tuning(*tiling)
            sol.set_schedule(sched)
            assert sol.schedule() == sched
            f0 = workloads.seeded_f(p, sol.grids(), nodes, noise=0.3, seed=9)
            sol.set_nodes(nodes, n_active)
            sol.upload_f(f0)
            for _ in range(2):
                sol.step()
            out.append(sol.download_f())
        finally:
            sol.close()
    return out, n_active


@pytest.mark.parametrize("tiling", [(512, 12, 2, 1), (256, 12, 2, 0)])
def test_mixed_precision_schedules_same_bits(tiling):
    """FP32 storage (precision mode 1): the persistent slab launch, its
    unbatched form and one launch per pass run the same float arithmetic
    per chunk, so two steps give the same bits, on one-chunk and two-chunk
    streams, thinned and ragged, one of them needing 147,956 streaming and
    55,986 focusing subcycles per substep (the oracle's counts for these
    nodes)."""
    out, n_active = mixed_precision_schedules(tiling)
    for s, n in enumerate(n_active):
        np.testing.assert_array_equal(out[1][s, :n], out[0][s, :n])
        np.testing.assert_array_equal(out[2][s, :n], out[0][s, :n])


@pytest.mark.skip(reason="these steps (FP32 storage, 8 subcycles per pass, one or three nodes per lane, on "
                         "the nodes above whose stream needs 147,956 subcycles per substep) ran past the "
                         "test's time limit on the GPU and the cause is not found; the library refuses "
                         "one node per lane and halos below 12 with FP32 storage")
@pytest.mark.parametrize("tiling", [(256, 8, 1, 0), (192, 8, 3, 0)])
def test_mixed_precision_small_halo_switched_off(tiling):
    out, n_active = mixed_precision_schedules(tiling)
    for s, n in enumerate(n_active):
        np.testing.assert_array_equal(out[1][s, :n], out[0][s, :n])


def test_fp32_refuses_one_node_per_lane():
    """One node per lane and halos below 12 are switched off with FP32
    storage: ep_set_tuning refuses them in precision modes 1 and 3 and
    accepts them again in FP64."""
    p = params(numEpSteps=1)
    nodes = workloads.parker_nodes(1, 1, 100, seed=66)
    dims = workloads.dims_of(nodes, 6, 8)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_tuning(256, 16, 1, 1)
        sol.set_precision(1)  # moves the tiling to two nodes per lane
        for mode in (1, 3):
            sol.set_precision(mode)
            for tiling in ((256, 16, 1, 0), (512, 8, 2, 0)):
                with pytest.raises(runtime.EpError) as ei:
                    sol.set_tuning(*tiling)
                assert ei.value.code == abi.EP_ERR_UNSUPPORTED
        sol.set_precision(1)
        sol.set_tuning(512, 12, 2, 0)
        sol.set_precision(0)
        sol.set_tuning(256, 16, 1, 1)
        sol.set_tuning(192, 8, 3, 0)
    finally:
        sol.close()


@pytest.mark.parametrize("tiling", [(256, 16, 1, 1), (512, 32, 2, 1), (384, 24, 2, 1), (256, 16, 1, 0),
                                    (576, 24, 3, 1)])
def test_fused_focusing_mixed_thinning(tiling):
    """Streams without thinned nodes are focused by their last streaming
    pass, thinned ones (filled after the pass) by the focusing kernel; a
    cube with both kinds matches the oracle over several steps."""
    p = params(numEpSteps=3)
    a = workloads.parker_nodes(1, 1, 150, seed=64, perturb=0.2, jitter=0.2, thin_frac=0.15)
    nodes = workloads.parker_nodes(1, 1, 150, seed=64, perturb=0.2, jitter=0.2)
    for k in nodes:
        nodes[k][0::2] = a[k][0::2]
    fg, fo, gd, od, _, _ = run_both(p, nodes, 12, 30, steps=2, tiling=tiling, precision=2)
    assert_diag_equal(gd, od)
    assert max_rel(fg, fo) <= RTOL


def test_two_nodes_per_lane_nan_location():
    """A NaN on either node of a lane's pair is reported at the cell the
    reference panics on."""
    p = params()
    nodes = workloads.parker_nodes(1, 1, 64, seed=9, perturb=0.2)
    g = oracle.build_grids(p, 12, 8)
    for node in (17, 18):
        f0 = workloads.seeded_f(p, g, nodes, noise=0.2, seed=1)
        f0[3, node, 5, 2] = np.nan
        fg, fo, gd, od, _, _ = run_both(p, nodes, 12, 8, f0=f0, tiling=(256, 8, 2))
        (rg, dg), (ro, do) = gd[0], od[0]
        assert rg == ro == abi.EP_ERR_NAN
        for k in ("err_stream", "err_node", "err_energy", "err_mu"):
            assert dg[k] == do[k], (node, k, dg[k], do[k])


def test_autotune_leaves_state_and_parity():
    """The tile-size search times passes without changing f; the tiling it
    picks gives results within the parity bar."""
    p = params()
    nodes = workloads.parker_nodes(1, 1, 300, seed=22, perturb=0.2, jitter=0.2)
    dims = workloads.dims_of(nodes, 16, 12)
    sol = runtime.EpSolver(dims, p)
    try:
        f0 = workloads.seeded_f(p, sol.grids(), nodes, noise=0.3, seed=6)
        sol.set_nodes(nodes)
        sol.upload_f(f0)
        block, halo, r, fz, ms = sol.autotune()
        assert r in (1, 2, 3) and block // r in (128, 192, 256, 320, 384, 512) and block % r == 0
        assert fz in (0, 1)
        assert 1 <= halo and 4 * halo <= block and ms > 0
        np.testing.assert_array_equal(sol.download_f(), f0)
        _, dg = sol.step()
        fg = sol.download_f()
    finally:
        sol.close()
    _, do, fo, _, _, _ = oracle.update_ep(dims, p, nodes, f0)
    assert_diag_equal([(0, dg)], [(0, do)])
    assert max_rel(fg, fo) <= RTOL


@pytest.mark.parametrize("tiling", [(256, 32, 1, 0), (256, 32, 1, 1), (512, 32, 2, 1), (512, 32, 2, 0)])
def test_heavy_focusing_nodes_parity(tiling):
    """Inner nodes with steep dlnB/ds need hundreds of focusing subcycles
    and take the one-lane-per-energy kernel; the rest are focused by the
    one-lane-per-node kernel or, fused, by the last streaming pass.  All
    must match the oracle."""
    p = params(rScale=0.00465047, numEpSteps=2)
    nodes = workloads.parker_nodes(1, 1, 300, rscale=0.00465047, seed=23, perturb=0.1, smooth=20,
                                   jitter=0.02)
    fg, fo, gd, od, _, _ = run_both(p, nodes, 20, 30, tiling=tiling)
    assert od[0][1]["max_subcycles_focusing"] >= 64  # heavy path taken
    assert_diag_equal(gd, od)
    assert max_rel(fg, fo) <= RTOL


def _ripple(a, n_active, head):
    """Shift every stream's node axis one node outward (numpy restatement of
    the reference's memcpy ripple); node 0 takes `head`."""
    out = a.copy()
    for s in range(a.shape[0]):
        n_new = min(n_active[s] + 1, a.shape[1])
        out[s, 1:n_new] = a[s, 0:n_new - 1]
        out[s, 0] = head[s]
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_spawn_head_moves_f_like_the_ripple(precision):
    """140 shell spawns with given spawn distributions: the device's f
    (head slot moved back per spawn, rows copied to the far end when the
    head reaches slot 0) equals the reference's ripple of every node one
    shell outward, bit for bit, for filling and full streams, in FP64 and
    FP32 storage; an EP step afterwards still matches the oracle."""
    p = params(numEpSteps=1)
    E, M = 5, 4
    nodes = workloads.parker_nodes(1, 1, 70, seed=35, jitter=0.1)
    S, N = nodes["x"].shape
    n_active = np.array([70, 40, 10, 69, 70, 2], np.int32)
    rng = np.random.default_rng(36)
    f = rng.uniform(0.5, 2.0, (S, N, E, M))
    dims = workloads.dims_of(nodes, E, M)
    sol = runtime.EpSolver(dims, p)
    cast = (lambda a: a.astype(np.float32).astype(np.float64)) if precision else (lambda a: a)
    try:
        sol.set_nodes(nodes, n_active)
        sol.set_precision(precision)
        sol.upload_f(f)
        fo, na = cast(f), n_active.copy()
        # spawn positions chosen so that the last 70 spawns lay out a stream
        final = workloads.parker_nodes(1, 1, N, seed=37, jitter=0.1)
        for j in range(140):
            sp = rng.uniform(0.5, 2.0, (S, E, M))
            k = min(139 - j, N - 1)
            sol.spawn_shell(*[final[q][:, k].copy() for q in ("x", "y", "z", "rmag")], f=sp)
            fo = _ripple(fo, na, cast(sp))
            na = np.minimum(na + 1, N)
            if j % 35 == 34:
                fg = sol.download_f()
                mask = active_mask(fo.shape[:2], na)
                np.testing.assert_array_equal(fg[mask], fo[mask], err_msg="spawn %d" % j)
        assert np.array_equal(sol.positions()["n_active"], na)
        if precision == 0:
            # the EP update reads the moved rows: one step against the oracle
            full = {k: _ripple_n(nodes[k], n_active, 140) for k in nodes}
            full.update({q: final[q] for q in ("x", "y", "z", "rmag")})
            rc, _ = sol.step(raise_on_error=False)
            fg = sol.download_f()
            rco, _, fe, _, _, _ = oracle.update_ep(dims, p, full, fo, n_active=na)
            assert rc == rco == 0
            assert max_rel(fg, fe, na) <= RTOL
    finally:
        sol.close()


def _ripple_n(a, n_active, times):
    """`times` ripples of a node array whose new node 0 repeats node 0."""
    na = np.array(n_active)
    for _ in range(times):
        a = _ripple(a, na, a[:, 0])
        na = np.minimum(na + 1, a.shape[1])
    return a


@pytest.mark.parametrize("steps", [3, 130])
def test_device_node_pipeline_parity(steps):
    """Global steps with node motion on the device -- spawn/ripple, move,
    updateMhd from the device-held state, EP update -- against the oracle
    doing the same steps (or_move_nodes is the function the oracle's
    analytic-wind pipeline uses, pinned by the wind digests).  130 steps
    run the f rows' head down to slot 0 and through the copy back."""
    from eprem_amd import constants as K, mas_synth
    rs = 0.00465047
    p = params(rScale=rs, numEpSteps=2)
    E, M = 10, 8
    nodes = workloads.parker_nodes(1, 1, 70, tdel_days=1.0 / 64, rscale=rs, seed=31, jitter=0.1)
    S, N = nodes["x"].shape
    n_active = np.full(S, N, np.int32)
    n_active[::2] = N - 3  # some streams still filling up
    prev = mas_synth.prev_from_nodes(nodes, rs)
    kw = dict(mhdUs=400e5 / K.C, mhdNsAu=8.3, mhdBsAu=2.5e-5 / K.MHD_B_NORM, omegaSun=0.0014)
    tdel = p.tDel
    g = oracle.build_grids(p, E, M)
    f = workloads.seeded_f(p, g, nodes, noise=0.2, seed=8)
    spawn_r = nodes["rmag"][:, 0].copy()
    zen0 = np.arccos(nodes["z"][:, 0] / spawn_r)
    azi0 = np.arctan2(nodes["y"][:, 0], nodes["x"][:, 0])

    def spawn_pos(j):
        azi = azi0 + 0.01 * (j + 1)
        return (spawn_r * np.sin(zen0) * np.cos(azi), spawn_r * np.sin(zen0) * np.sin(azi),
                spawn_r * np.cos(zen0), spawn_r.copy())

    dims = workloads.dims_of(nodes, E, M)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_positions(nodes, n_active)
        sol.update_mhd(prev, mode=0, **kw)
        sol.upload_f(f)
        for j in range(steps):
            sol.spawn_shell(*spawn_pos(j))
            sol.move_nodes(tdel)
            sol.update_mhd(None, mode=0, **kw)
            sol.step()
        pos_g = sol.positions()
        fg = sol.download_f()
    finally:
        sol.close()

    # oracle side
    o_nodes = {k: nodes[k].copy() for k in ("x", "y", "z", "rmag")}
    state = oracle.update_mhd(dict(o_nodes), prev, 0, rs, **kw)
    na = n_active.copy()
    fo = f.copy()
    for j in range(steps):
        sx, sy, sz, sr = spawn_pos(j)
        for k, h in zip(("x", "y", "z", "rmag"), (sx, sy, sz, sr)):
            o_nodes[k] = _ripple(o_nodes[k], na, h)
        for k in ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density"):
            state[k] = _ripple(state[k], na, state[k][:, 0])
        seed0 = np.array([[oracle.seed(p, g["egrid"][e], sr[s]) for e in range(E)] for s in range(S)])
        fo = _ripple(fo, na, np.repeat(seed0[:, :, None], M, axis=2))
        na = np.minimum(na + 1, N)
        old = oracle.move_nodes(o_nodes, state, tdel, rs, first=1)
        pv = dict(state)
        pv.update(old)
        state = oracle.update_mhd(dict(o_nodes), pv, 0, rs, **kw)
        full = dict(o_nodes)
        full.update({k: state[k] for k in ("bmag", "dlnb", "dlnn", "dupar")})
        rc, _, fo, _, _, _ = oracle.update_ep(dims, p, full, fo, n_active=na)
        assert rc == 0
    np.testing.assert_array_equal(pos_g["n_active"], na)
    act = np.arange(N)[None, :] < na[:, None]
    for k in ("x", "y", "z", "rmag"):
        np.testing.assert_allclose(pos_g[k][act], o_nodes[k][act], rtol=1e-13, atol=1e-11, err_msg=k)
    assert max_rel(fg, fo, na) <= RTOL


def test_device_mas_node_pipeline_parity():
    """Global steps of a MAS-coupled run on the device -- spawn/ripple,
    masMoveNodes (RK4 through the two MAS levels inside mas_radial_max,
    masWind of the node's own state outside), updateMhd from MAS, EP
    update -- against the oracle doing the same steps.  The nodes span
    1-54 Rs, so both branches of masGetNode act at every RK4 stage."""
    from eprem_amd import constants as K, mas_synth
    rs = 0.00465047
    p = params(rScale=rs, numEpSteps=2)
    E, M = 8, 6
    nodes = workloads.parker_nodes(1, 1, 70, tdel_days=1.0 / 64, rscale=rs, seed=33, jitter=0.1)
    S, N = nodes["x"].shape
    n_active = np.full(S, N, np.int32)
    n_active[1::2] = N - 4
    prev = mas_synth.prev_from_nodes(nodes, rs)
    r, t, ph = mas_synth.meshes(nr=120, nt=90, nphi=180)
    lv0 = mas_synth.level(r, t, ph, cme_r=8.0, seed=5)
    lv1 = mas_synth.level(r, t, ph, cme_r=12.0, seed=5)
    mas = dict(s_cor=0.3, phi_offset=0.2, mas_radial_max=30.0, mas_vmin=200e5)
    inside = nodes["rmag"] * rs <= mas["mas_radial_max"] * K.RSAU
    assert 0.2 < inside.mean() < 0.8
    tdel = p.tDel
    g = oracle.build_grids(p, E, M)
    f = workloads.seeded_f(p, g, nodes, noise=0.2, seed=9)
    spawn_r = nodes["rmag"][:, 0].copy()
    zen0 = np.arccos(nodes["z"][:, 0] / spawn_r)
    azi0 = np.arctan2(nodes["y"][:, 0], nodes["x"][:, 0])
    steps = 3

    def spawn_pos(j):
        azi = azi0 + 0.01 * (j + 1)
        return (spawn_r * np.sin(zen0) * np.cos(azi), spawn_r * np.sin(zen0) * np.sin(azi),
                spawn_r * np.cos(zen0), spawn_r.copy())

    dims = workloads.dims_of(nodes, E, M)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.mas_set_level(0, *lv0)
        sol.mas_set_level(1, *lv1)
        sol.set_positions(nodes, n_active)
        sol.update_mhd(prev, mode=1, **mas)
        sol.upload_f(f)
        for j in range(steps):
            sol.spawn_shell(*spawn_pos(j))
            sol.move_nodes_mas(tdel, **mas)
            sol.update_mhd(None, mode=1, **mas)
            rc, _ = sol.step(raise_on_error=False)
            assert rc == 0
        pos_g = sol.positions()
        fg = sol.download_f()
    finally:
        sol.close()

    o_nodes = {k: nodes[k].copy() for k in ("x", "y", "z", "rmag")}
    lv = dict(level0=lv0, level1=lv1, **mas)
    state = oracle.update_mhd(dict(o_nodes), prev, 1, rs, **lv)
    rold = {k: prev[k].copy() for k in ("rold_x", "rold_y", "rold_z")}
    na = n_active.copy()
    fo = f.copy()
    for j in range(steps):
        sx, sy, sz, sr = spawn_pos(j)
        for k, h in zip(("x", "y", "z", "rmag"), (sx, sy, sz, sr)):
            o_nodes[k] = _ripple(o_nodes[k], na, h)
        for k, h in zip(("rold_x", "rold_y", "rold_z"), (sx, sy, sz)):
            rold[k] = _ripple(rold[k], na, h)
        for k in oracle.MHD_NAMES:
            state[k] = _ripple(state[k], na, state[k][:, 0])
        seed0 = np.array([[oracle.seed(p, g["egrid"][e], sr[s]) for e in range(E)] for s in range(S)])
        fo = _ripple(fo, na, np.repeat(seed0[:, :, None], M, axis=2))
        na = np.minimum(na + 1, N)
        st = {k: state[k] for k in oracle.MHD_NAMES}
        st.update(rold)
        rold = oracle.move_nodes_mas(o_nodes, st, tdel, rs, lv0, lv1, mas["s_cor"], mas["phi_offset"],
                                     mas["mas_radial_max"], mas["mas_vmin"], n_active=na, first=1)
        pv = {k: state[k] for k in oracle.MHD_NAMES}
        pv.update(rold)
        state = oracle.update_mhd(dict(o_nodes), pv, 1, rs, **lv)
        full = dict(o_nodes)
        full.update({k: state[k] for k in ("bmag", "dlnb", "dlnn", "dupar")})
        rc, _, fo, _, _, _ = oracle.update_ep(dims, p, full, fo, n_active=na)
        assert rc == 0
    np.testing.assert_array_equal(pos_g["n_active"], na)
    act = np.arange(N)[None, :] < na[:, None]
    # the nodes moved, and differently from a straight radial step
    assert np.max(np.abs(o_nodes["rmag"][act] - _ripple(nodes["rmag"], n_active, spawn_r)[act])) > 0
    for k in ("x", "y", "z", "rmag"):
        np.testing.assert_allclose(pos_g[k][act], o_nodes[k][act], rtol=1e-13, atol=1e-11, err_msg=k)
    assert max_rel(fg, fo, na) <= RTOL


def test_mixed_precision_error_bounds():
    """FP32 streaming/focusing with FP64 adiabatic change against the FP64
    oracle on a 35-decade power-law distribution: relative error of f and
    of the omnidirectional flux at three observers stays far inside the 1%
    gate, and the FP64 mode is restored exactly by switching back."""
    p = params(rScale=0.00465047, numEpSteps=3)
    nodes = workloads.parker_nodes(1, 1, 200, rscale=0.00465047, seed=41, perturb=0.1, smooth=20,
                                   jitter=0.02)
    E, M = 20, 30
    dims = workloads.dims_of(nodes, E, M)
    g = oracle.build_grids(p, E, M)
    f0 = workloads.power_law_f(g, nodes, M, seed=3)
    assert f0.max() > 1e4 and f0.min() < 1e-20
    r_au = np.array([0.3, 0.5, 0.8])
    th = np.array([1.0, 1.6, 2.2])
    ph = np.array([0.5, 2.0, 4.0])
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_nodes(nodes)
        sol.upload_f(f0)
        sol.set_precision(True)
        for _ in range(2):
            sol.step()
        f32 = sol.download_f()
        _, flux32 = sol.observe(r_au, th, ph)
        sol.set_precision(False)
        sol.upload_f(f0)
        sol.step()
        f64_1 = sol.download_f()
    finally:
        sol.close()
    fo = f0.copy()
    for i in range(2):
        rc, _, fo, _, _, _ = oracle.update_ep(dims, p, nodes, fo)
        assert rc == 0
        if i == 0:
            assert max_rel(f64_1, fo) <= RTOL  # FP64 mode after a precision round trip
    _, flux64 = oracle.observers(dims, nodes, fo, g["egrid"], p.rScale, r_au, th, ph, 3.0)
    # float storage floors at FLT_MIN; the error is stated over the range the
    # workload spans (f >= 1e-30)
    live = fo >= 1e-30
    assert live.mean() > 0.9
    rel_f = np.abs(f32[live] - fo[live]) / fo[live]
    assert np.median(rel_f) < 1e-5 and rel_f.max() < 1e-3, (np.median(rel_f), rel_f.max())
    rel_flux = np.abs(flux32 - flux64) / flux64
    assert rel_flux.max() < 1e-4, rel_flux.max()


def test_log_scaled_storage():
    """Precision mode 3 (mode 1 with the float storage holding ln f): the
    same operators on f decoded from FP32 ln f.  Against the FP64 oracle on
    the 35-decade power law, f and the observer flux stay within the bounds
    FP32 ln f can hold (2^-24 |ln f| per rounding, ~4e-6 relative at
    |ln f| = 70, several roundings per step); upload and download keep
    values far below FLT_MIN; switching modes converts through FP64."""
    p = params(rScale=0.00465047, numEpSteps=3)
    nodes = workloads.parker_nodes(1, 1, 200, rscale=0.00465047, seed=41, perturb=0.1, smooth=20,
                                   jitter=0.02)
    E, M = 20, 30
    dims = workloads.dims_of(nodes, E, M)
    g = oracle.build_grids(p, E, M)
    f0 = workloads.power_law_f(g, nodes, M, seed=3)
    r_au = np.array([0.3, 0.5, 0.8])
    th = np.array([1.0, 1.6, 2.2])
    ph = np.array([0.5, 2.0, 4.0])
    tiny = f0.copy()
    tiny[0, 5, 3, :] = 1e-300
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_nodes(nodes)
        sol.set_precision(3)
        sol.upload_f(tiny)
        back = sol.download_f()
        sol.set_precision(1)          # ln f -> f (float), through FP64
        lin = sol.download_f()
        sol.set_precision(3)
        # mode 3 has the slab launch with two nodes per lane, focusing in
        # its own kernel: other tilings and the per-pass schedule are refused
        assert sol.schedule() in (1, 2)
        with pytest.raises(runtime.EpError):
            sol.set_schedule(0)
        with pytest.raises(runtime.EpError):
            sol.set_tuning(256, 16, lane_nodes=1, fuse_focus=0)
        with pytest.raises(runtime.EpError):
            sol.set_tuning(512, 16, lane_nodes=2, fuse_focus=1)
        # the tile-size search keeps to what mode 3 has kernels for
        _, _, lane_nodes, fuse, _ = sol.autotune()
        assert lane_nodes == 2 and fuse == 0 and sol.schedule() in (1, 2)
        sol.upload_f(f0)
        for _ in range(2):
            sol.step()
        f3 = sol.download_f()
        _, flux3 = sol.observe(r_au, th, ph)
        sol.set_precision(0)
        back64 = sol.download_f()
    finally:
        sol.close()
    np.testing.assert_allclose(back, tiny, rtol=1e-4)  # 1e-300 survives the log storage
    keep = tiny > 1e-30  # linear float storage cannot hold the 1e-300 row
    np.testing.assert_allclose(lin[keep], tiny[keep], rtol=1e-5)
    assert np.all(lin[~keep] == 0.0)
    np.testing.assert_allclose(back64, f3, rtol=0, atol=0)
    fo = f0.copy()
    for _ in range(2):
        rc, _, fo, _, _, _ = oracle.update_ep(dims, p, nodes, fo)
        assert rc == 0
    _, flux64 = oracle.observers(dims, nodes, fo, g["egrid"], p.rScale, r_au, th, ph, 3.0)
    live = fo >= 1e-30
    rel_f = np.abs(f3[live] - fo[live]) / fo[live]
    rel_flux = np.abs(flux3 - flux64) / flux64
    print("log-scaled storage vs the FP64 oracle: f median %.2e max %.2e, flux max %.2e"
          % (np.median(rel_f), rel_f.max(), rel_flux.max()))
    assert np.median(rel_f) < 1e-4 and rel_f.max() < 1e-2, (np.median(rel_f), rel_f.max())
    assert rel_flux.max() < 1e-3, rel_flux.max()


def test_mixed_precision_interpolation():
    """In the mixed-precision form the MAS trilinear interpolation runs in
    FP32 (search fractions, stencil, time blend; unit conversions in FP64):
    every output within a few FP32 roundings of the oracle's FP64
    masTriLinear, scaled by the output's largest magnitude (a mean-free
    scale: B_theta and B_phi pass through zero), and the FP64 interpolation
    is back to 1e-13 after switching to precision mode 0."""
    from eprem_amd import mas_synth
    r, t, ph = mas_synth.meshes(nr=40, nt=30, nphi=50, rmin=1.0, rmax=30.0)
    lv0 = mas_synth.level(r, t, ph, cme_r=8.0, seed=3)
    lv1 = mas_synth.level(r, t, ph, cme_r=12.0, seed=3)
    nodes = workloads.parker_nodes(2, 2, 80, tdel_days=1.0 / 512, rscale=0.00465047, seed=13, jitter=0.3)
    p = params(rScale=0.00465047)
    dims = workloads.dims_of(nodes, 4, 4)
    prev = mas_synth.prev_from_nodes(nodes, p.rScale)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_nodes(nodes)
        sol.mas_set_level(0, *lv0)
        sol.mas_set_level(1, *lv1)
        sol.set_precision(1)
        o32 = sol.interp_mhd(0.37, 0.8, 30.0, 50.0e5, prev)
        sol.set_precision(0)
        o64 = sol.interp_mhd(0.37, 0.8, 30.0, 50.0e5, prev)
    finally:
        sol.close()
    ref = oracle.mas_interp(p.rScale, nodes["x"].ravel(), nodes["y"].ravel(), nodes["z"].ravel(),
                            nodes["rmag"].ravel(), lv0, lv1, 0.37, 0.8, 30.0, 50.0e5,
                            {k: v.ravel() for k, v in prev.items()})
    inside = (nodes["rmag"].ravel() * p.rScale) <= 30.0 * (6.96e10 / 1.495978707e13)
    assert inside.mean() > 0.5
    moved = 0
    for k in ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density"):
        a, b = o32[k].ravel()[inside], ref[k][inside]
        scale = np.abs(b).max()
        if scale == 0.0:  # the synthetic wind has no theta components and no V_phi
            np.testing.assert_array_equal(a, b, err_msg=k)
        else:
            err = np.abs(a - b).max() / scale
            assert err < 1e-5, (k, err)
            moved += err > 0.0
        np.testing.assert_allclose(o64[k].ravel(), ref[k], rtol=1e-13, atol=0, err_msg=k)
    assert moved >= 3  # the FP32 arithmetic is really in use


@pytest.mark.parametrize("case", [
    dict(N=64, E=12, M=6, steps=1, tiling=None),
    dict(N=300, E=20, M=30, steps=3, tiling=(128, 8)),
    dict(N=200, E=16, M=7, steps=2, tiling=(512, 100)),
])
def test_fused_streaming_parity(case):
    """Precision mode 2 (fused multiply-adds in streaming and relaxation)
    is not bit-identical to the reference's operation order but stays
    within the parity bar, with identical subcycle counts."""
    p = params(rScale=0.00465047)
    nodes = workloads.parker_nodes(1, 1, case["N"], rscale=0.00465047, seed=51, perturb=0.1, smooth=20,
                                   jitter=0.05)
    fg, fo, gd, od, _, _ = run_both(p, nodes, case["E"], case["M"], steps=case["steps"],
                                    tiling=case["tiling"], precision=2)
    assert_diag_equal(gd, od)
    assert max_rel(fg, fo) <= RTOL


@pytest.mark.parametrize("rows,kp", [(2, 0.5), (3, 50.0)])
def test_shell_diffusion_parity(rows, kp):
    """DiffuseShellData on the device (gather in source order) against the
    oracle's scatter; kp = 50 drives most links to the THRESH cap."""
    from eprem_amd import constants as K, cube, mas_synth
    rs = 0.00465047
    p = params(rScale=rs, numEpSteps=2)
    nodes = workloads.parker_nodes(rows, rows, 40, rscale=rs, seed=71, jitter=0.1)
    S, N = nodes["x"].shape
    E, M = 6, 5
    prev = mas_synth.prev_from_nodes(nodes, rs)
    kw = dict(mhdUs=400e5 / K.C, mhdNsAu=8.3, mhdBsAu=2.5e-5 / K.MHD_B_NORM, omegaSun=0.0014)
    nb = cube.neighbours(rows, rows)
    dims = workloads.dims_of(nodes, E, M)
    g = oracle.build_grids(p, E, M)
    f0 = workloads.seeded_f(p, g, nodes, noise=0.4, seed=12)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_positions(nodes)
        out = sol.update_mhd(prev, mode=0, **kw)
        sol.upload_f(f0)
        sol.shell_operators(nb, kp)
        rc, _ = sol.step(abi.EP_OP_SHELL)
        fg = sol.download_f()
    finally:
        sol.close()
    assert rc == 0
    fo = f0.copy()
    for _ in range(p.numEpSteps):
        rc, fo, _ = oracle.shell_diffuse(p, nodes, out, fo, nb, g, p.tDel / p.numEpSteps, kp)
        assert rc == 0
    assert not np.array_equal(fo, f0)
    assert max_rel(fg, fo) <= 1e-12


def _drift_oracle(p, nodes, out, kw, f0, nb, g, kp, drift):
    """Oracle reference for the cross-stream operators: per substep
    DiffuseShellData (kp > 0), then DriftShellData with curlBoverB2 of the
    analytic field at every node."""
    S, N = nodes["x"].shape
    curl = np.zeros((S, N, 3))
    if drift:
        for s in range(S):
            for k in range(1, N):
                curl[s, k] = oracle.curl_b_over_b2(p.rScale, kw["mhdBsAu"], kw["omegaSun"], nodes["x"][s, k],
                                                   nodes["y"][s, k], nodes["z"][s, k], out["vr"][s, k])
    fo = f0.copy()
    dt = p.tDel / p.numEpSteps
    for _ in range(p.numEpSteps):
        if kp > 0:
            rc, fo, _ = oracle.shell_diffuse(p, nodes, out, fo, nb, g, dt, kp)
            assert rc == 0
        if drift:
            rc, fo, _ = oracle.shell_drift(p, nodes, curl, fo, nb, g, dt)
            assert rc == 0
    return fo


@pytest.mark.parametrize("kp,charge", [(0.0, 1.0), (0.5, 1.0), (0.0, -1.0)])
def test_shell_drift_parity(kp, charge):
    """DriftShellData (gradient/curvature drift from curlBoverB2 of the
    analytic Parker field) on the device, alone and after diffusion,
    against the oracle's scatter; charge -1 reverses the drift."""
    from eprem_amd import constants as K, cube, mas_synth
    rs = 0.00465047
    p = params(rScale=rs, numEpSteps=2, charge=[charge])
    nodes = workloads.parker_nodes(3, 3, 40, rscale=rs, seed=73, jitter=0.1)
    E, M = 6, 5
    prev = mas_synth.prev_from_nodes(nodes, rs)
    kw = dict(mhdUs=400e5 / K.C, mhdNsAu=8.3, mhdBsAu=2.5e-5 / K.MHD_B_NORM, omegaSun=0.0014)
    nb = cube.neighbours(3, 3)
    dims = workloads.dims_of(nodes, E, M)
    g = oracle.build_grids(p, E, M)
    f0 = workloads.seeded_f(p, g, nodes, noise=0.4, seed=13)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_positions(nodes)
        out = sol.update_mhd(prev, mode=0, **kw)
        sol.upload_f(f0)
        sol.shell_operators(nb, kp, drift=True)
        rc, _ = sol.step(abi.EP_OP_SHELL | abi.EP_OP_DRIFT)
        fg = sol.download_f()
    finally:
        sol.close()
    assert rc == 0
    fo = _drift_oracle(p, nodes, out, kw, f0, nb, g, kp, True)
    assert max_rel(fo, f0) > 1e-6
    assert max_rel(fg, fo) <= 1e-12

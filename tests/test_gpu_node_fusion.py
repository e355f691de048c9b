"""The one-pass node kernel (AdiabaticFocusing then AdiabaticChange of a
node, src/energeticParticles.c:1533-1803 and :805-1077, with f read and
written once) against the two-kernel form and the FP64 oracle.

The one-pass kernel repeats the two kernels' arithmetic operation for
operation, so the two forms must agree bit for bit in every precision mode;
against the oracle the usual 1e-10 bar applies.  Cases cover odd and even
mu counts (the mu row is split over a lane pair), heavy-focusing nodes
(focused beforehand by the one-lane-per-energy kernel), thinned and ragged
streams, focusing fused into the streaming passes, and precision modes
0, 1 and 2.
"""
import numpy as np
import pytest

import oracle
from eprem_amd import abi, config, runtime, workloads

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def params(**kw):
    base = dict(numEpSteps=8, tDel=1.0 / 64, eMin=0.01, eMax=1000.0, lamo=0.1, mfpRadialPower=1.0,
                boundaryFunctEcutoff=1.0e30, rScale=0.00465047)
    base.update(kw)
    return config.default_params(**base)


def run_gpu(p, nodes, E, M, f0, fused, n_active=None, tiling=None, precision=0, steps=1,
            op_flags=abi.EP_OP_ALL):
    dims = workloads.dims_of(nodes, E, M)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_node_fusion(fused)
        if tiling:
            sol.set_tuning(*tiling)
        sol.set_nodes(nodes, n_active)
        sol.upload_f(f0)
        if precision:
            sol.set_precision(precision)
        diags = []
        for _ in range(steps):
            rc, dg = sol.step(op_flags, raise_on_error=False)
            diags.append((rc, dg))
            if rc:
                break
        if precision == 1:
            sol.set_precision(0)
        return sol.download_f(), diags
    finally:
        sol.close()


CASES = [
    dict(N=200, E=20, M=30, seed=61),
    dict(N=130, E=12, M=7, seed=62),
    dict(N=100, E=9, M=5, seed=63),
    dict(N=90, E=64, M=32, seed=64),
    dict(N=70, E=2, M=2, seed=65),
    dict(N=300, E=16, M=31, seed=66),
    # four lanes per node from M = 12: cell M-1 first in its group (13),
    # second (17), third (12), fourth (25)
    dict(N=150, E=12, M=13, seed=70),
    dict(N=140, E=10, M=17, seed=71),
    dict(N=130, E=9, M=12, seed=72),
    dict(N=110, E=11, M=25, seed=73),
]


@pytest.mark.parametrize("precision", [0, 2, 1])
@pytest.mark.parametrize("case", CASES)
def test_one_pass_equals_two_kernels(case, precision):
    p = params()
    nodes = workloads.parker_nodes(1, 1, case["N"], rscale=0.00465047, seed=case["seed"], perturb=0.1,
                                   smooth=20, jitter=0.05)
    g = oracle.build_grids(p, case["E"], case["M"])
    f0 = workloads.seeded_f(p, g, nodes, noise=0.3, seed=case["seed"])
    f1, d1 = run_gpu(p, nodes, case["E"], case["M"], f0, 1, precision=precision)
    f2, d2 = run_gpu(p, nodes, case["E"], case["M"], f0, 0, precision=precision)
    assert d1[0][0] == d2[0][0] == 0
    assert d1[0][1]["max_subcycles_change"] <= 1  # the one-pass kernel ran
    np.testing.assert_array_equal(f1, f2)
    if precision == 0:
        dims = workloads.dims_of(nodes, case["E"], case["M"])
        rc, do, fo, _, _, _ = oracle.update_ep(dims, p, nodes, f0)
        assert rc == 0
        assert float(np.max(np.abs(f1 - fo) / np.abs(fo))) <= RTOL


@pytest.mark.parametrize("tiling", [(256, 32, 1, 0), (256, 32, 1, 1), (512, 48, 2, 1), (512, 48, 2, 0)])
def test_one_pass_heavy_nodes_thinning_ragged(tiling):
    """Heavy-focusing inner nodes (hundreds of subcycles), thinned nodes,
    ragged active counts, focusing in the streaming passes or not, eight
    substeps: bit-identical to the two-kernel form and within the bar of
    the oracle."""
    p = params()
    nodes = workloads.parker_nodes(1, 1, 300, rscale=0.00465047, seed=67, perturb=0.1, smooth=20,
                                   jitter=0.02, thin_frac=0.05)
    n_active = np.array([300, 41, 255, 129, 3, 200], np.int32)
    E, M = 20, 30
    g = oracle.build_grids(p, E, M)
    f0 = workloads.seeded_f(p, g, nodes, noise=0.3, seed=5)
    f1, d1 = run_gpu(p, nodes, E, M, f0, 1, n_active=n_active, tiling=tiling)
    f2, d2 = run_gpu(p, nodes, E, M, f0, 0, n_active=n_active, tiling=tiling)
    np.testing.assert_array_equal(f1, f2)
    dims = workloads.dims_of(nodes, E, M)
    rc, do, fo, _, _, _ = oracle.update_ep(dims, p, nodes, f0, n_active=n_active)
    assert rc == 0 and d1[0][0] == 0
    assert do["max_subcycles_focusing"] >= 64 and do["max_subcycles_change"] <= 1
    for k in ("max_subcycles_change", "max_subcycles_focusing", "max_nsteps_stream"):
        assert d1[0][1][k] == do[k], k
    mask = np.zeros(f1.shape[:2], bool)
    for s, n in enumerate(n_active):
        mask[s, :n] = True
    assert float(np.max(np.abs(f1[mask] - fo[mask]) / np.abs(fo[mask]))) <= RTOL


@pytest.mark.parametrize("flags", [abi.EP_OP_FOCUS | abi.EP_OP_CHANGE,
                                   abi.EP_OP_FOCUS | abi.EP_OP_CHANGE | abi.EP_OP_SEEDCHECK])
def test_one_pass_without_streaming(flags):
    """Focusing and change alone (streaming off), against the oracle."""
    p = params()
    nodes = workloads.parker_nodes(1, 1, 150, rscale=0.00465047, seed=68, perturb=0.1, smooth=20,
                                   jitter=0.05)
    E, M = 14, 11
    g = oracle.build_grids(p, E, M)
    f0 = workloads.seeded_f(p, g, nodes, noise=0.3, seed=8)
    f1, d1 = run_gpu(p, nodes, E, M, f0, 1, op_flags=flags)
    f2, _ = run_gpu(p, nodes, E, M, f0, 0, op_flags=flags)
    np.testing.assert_array_equal(f1, f2)
    dims = workloads.dims_of(nodes, E, M)
    rc, do, fo, _, _, _ = oracle.update_ep(dims, p, nodes, f0, op_flags=flags)
    assert rc == 0 and d1[0][0] == 0
    assert float(np.max(np.abs(f1 - fo) / np.abs(fo))) <= RTOL


@pytest.mark.parametrize("M", [11, 13])
@pytest.mark.parametrize("flags", [abi.EP_OP_FOCUS | abi.EP_OP_CHANGE, abi.EP_OP_CHANGE])
@pytest.mark.parametrize("where,value", [((2, 17, 5, 2), np.nan), ((0, 150, 0, 0), np.nan),
                                         ((4, 60, 13, 10), np.nan), ((1, 40, 7, 5), np.inf),
                                         ((3, 90, 0, 10), np.nan), ((5, 33, 13, 0), np.inf),
                                         ((2, 70, 3, -1), np.nan), ((1, 20, 9, -2), np.inf)])
def test_one_pass_nan_location(where, value, flags, M):
    """A NaN or Inf planted in f: the same error code and first-panic cell
    as the oracle (and as the two-kernel form).  The reference's interior
    fluxes pass a non-finite value on from either neighbour, its wall rows
    from the upwind cell only; cells on the mu and energy walls are among
    the planted ones."""
    p = params()
    nodes = workloads.parker_nodes(1, 1, 160, rscale=0.00465047, seed=69, perturb=0.1, smooth=20,
                                   jitter=0.05)
    E = 14
    g = oracle.build_grids(p, E, M)
    f0 = workloads.seeded_f(p, g, nodes, noise=0.2, seed=9)
    f0[where] = value  # (a negative mu index counts from the upper wall)
    _, d1 = run_gpu(p, nodes, E, M, f0, 1, op_flags=flags)
    _, d2 = run_gpu(p, nodes, E, M, f0, 0, op_flags=flags)
    dims = workloads.dims_of(nodes, E, M)
    ro, do, _, _, _, _ = oracle.update_ep(dims, p, nodes, f0, op_flags=flags)
    (r1, g1), (r2, g2) = d1[0], d2[0]
    assert ro in (abi.EP_ERR_NAN, abi.EP_ERR_INF)
    assert r1 == r2 == ro
    for k in ("err_stream", "err_node", "err_energy", "err_mu"):
        assert g1[k] == g2[k] == do[k], (k, g1[k], g2[k], do[k])

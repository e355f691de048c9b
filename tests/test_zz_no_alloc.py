"""ep_observe_noalloc (the observer entry point with fixed device scratch
and no allocation) against ep_observe, through the C ABI: identical
results on two calls in a row, the second with fewer observers and a
different state, and the size limit reported as an error before any
device work."""
import numpy as np
import pytest

import oracle
from eprem_amd import abi, config, runtime, workloads

pytestmark = pytest.mark.gpu


def test_noalloc_observe_equals_observe():
    p = config.default_params(numEpSteps=2, tDel=1.0 / 64, eMin=0.01, eMax=1000.0, lamo=0.1,
                              mfpRadialPower=1.0, boundaryFunctEcutoff=1.0e30)
    nodes = workloads.parker_nodes(1, 1, 200, seed=71, perturb=0.1, smooth=10)
    dims = workloads.dims_of(nodes, 12, 8)
    sol = runtime.EpSolver(dims, p)
    try:
        f0 = workloads.seeded_f(p, sol.grids(), nodes, noise=0.3, seed=72)
        sol.set_nodes(nodes)
        sol.upload_f(f0)
        calls = [
            (np.array([0.3, 0.5, 1.0, 0.7, 2.0]), np.array([1.2, 1.571, 2.0, 0.4, 2.9]),
             np.array([0.5, 2.5, 4.5, -1.0, 6.5])),
            (np.array([0.45, 1.3]), np.array([0.9, 2.2]), np.array([3.3, 0.1])),
        ]
        for i, (r, t, ph) in enumerate(calls):
            if i == 1:
                sol.step()  # a different state for the second call
            ref = sol.observe(r, t, ph, idw_p=3.0)
            got = sol.observe(r, t, ph, idw_p=3.0, noalloc=True)
            np.testing.assert_array_equal(got[0], ref[0])
            np.testing.assert_array_equal(got[1], ref[1])
            assert np.all(got[1] > 0)
        # and the first state's result against the oracle
        g = oracle.build_grids(p, 12, 8)
        r, t, ph = calls[0]
        sol.upload_f(f0)
        got = sol.observe(r, t, ph, idw_p=3.0, noalloc=True)
        d_o, f_o = oracle.observers(dims, nodes, f0, g["egrid"], p.rScale, r, t, ph, 3.0)
        np.testing.assert_allclose(got[0], d_o, rtol=1e-13, atol=0)
        np.testing.assert_allclose(got[1], f_o, rtol=1e-13, atol=0)
    finally:
        sol.close()


def test_noalloc_observe_rejects_oversized_requests():
    """256 observers of 24 streams x 64 x 32 cells need 101 MB of
    projections, more than the 32 MiB scratch: the call fails cleanly before
    any device work, and ep_observe (which allocates) still serves it."""
    p = config.default_params(numEpSteps=1, tDel=1.0 / 64, eMin=0.01, eMax=1000.0)
    nodes = workloads.parker_nodes(2, 2, 40, seed=73)
    dims = workloads.dims_of(nodes, 64, 32)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_nodes(nodes)
        sol.init_seed_population()
        n = 256
        r, t, ph = np.full(n, 0.2), np.full(n, 1.0), np.linspace(0.1, 6.0, n)
        with pytest.raises(runtime.EpError) as ei:
            sol.observe(r, t, ph, noalloc=True)
        assert ei.value.code == abi.EP_ERR_INVALID_ARG
        assert "scratch" in str(ei.value)
        d, f = sol.observe(r, t, ph)
        assert np.all(np.isfinite(f))
    finally:
        sol.close()

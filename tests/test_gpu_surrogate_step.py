"""The learned-surrogate form of ep_step (ep_set_surrogate).

* one fused network application (k_sur_apply) equals the large step
  times exp of the reference MLP forward (ep_mlp_forward) on the sample
  generator's inputs for the same (stream, energy, node), keeps node 0,
  and conserves sum_mu f at the last node (relaxation-only large step);
* the residual's plain check (a one-wave plain solve over each sampled
  node's 2K + 1 dependence window) reproduces the full plain pass at the
  sampled nodes to rounding;
* with a zero tolerance every stream falls back and the step is the plain
  step bit for bit; with no tolerance nothing falls back, graph and eager
  launches agree bit for bit, and the slab counters add up.
"""
import numpy as np
import pytest

from eprem_amd import config, runtime, workloads

pytestmark = pytest.mark.gpu

RS = 0.00465047


def _params(**kw):
    base = dict(rScale=RS, tDel=1.0 / 64, numEpSteps=4, eMin=0.01, eMax=1000.0, lamo=0.1,
                rigidityPower=1.0 / 3.0, mfpRadialPower=1.0, boundaryFunctEcutoff=1.0e30)
    base.update(kw)
    return config.default_params(**base)


def _solver(M=30, E=16, rows=2, cols=2, N=300, seed=71):
    p = _params()
    nodes = workloads.parker_nodes(rows, cols, N, rscale=RS, seed=seed, perturb=0.1, smooth=20, jitter=0.05)
    dims = workloads.dims_of(nodes, E, M)
    sol = runtime.EpSolver(dims, p)
    sol.set_nodes(nodes)
    sol.init_seed_population()
    sol.set_precision(2)
    sol.step()  # anisotropy from the isotropic seed
    return sol, dims


def _network(M, seed=5, scale=0.3):
    net = runtime.Mlp([3 * M + 6, 128, 128, 128, M], seed=seed)
    prm = net.params()
    net.set_params((prm * scale).astype(np.float32))
    return net


@pytest.mark.parametrize("M", [30, 12])
def test_apply_equals_network_forward(M):
    sol, dims = _solver(M=M)
    net = _network(M)
    try:
        K = 8
        sol.set_surrogate(K, 1e30)
        x, _, where = sol.surrogate_samples(K, 3000, seed=4, where=True)
        assert len(x) > 500
        f0 = sol.download_f()
        sol.surrogate_probe(0)
        f1 = sol.download_f()
        y = net.forward(x)
    finally:
        net.close()
        sol.close()
    s, e, k = where[:, 0], where[:, 1], where[:, 2]
    # the large step (exact relaxation over K dtProp) times exp of the
    # network's correction
    fk = f0[s, k, e, :]
    iso = fk.mean(axis=1, keepdims=True)
    qK = np.exp(-x[:, 3 * M + 1].astype(np.float64))[:, None]
    want = (iso + (fk - iso) * qK) * np.exp(y.astype(np.float64))
    got = f1[s, k, e, :]
    rel = np.abs(got - want) / want
    # the two paths round the same inputs to bf16 from FP32 values that
    # differ by ~1e-6 (FP64 log vs FP32 log of the mantissa), which moves an
    # input by one bf16 step now and then; the large step's q^K comes from
    # an FP32 input here
    assert np.quantile(rel, 0.95) < 2e-3, np.quantile(rel, [0.5, 0.95, 1.0])
    assert np.median(rel) < 2e-4
    assert rel.max() < 0.2
    # the network is not the identity on these rows
    assert np.median(np.abs(got - f0[s, k, e, :]) / f0[s, k, e, :]) > 1e-2
    # node 0 keeps its row; the last node's relaxation keeps sum_mu f
    qual = np.unique(np.stack([s, e], 1), axis=0)
    n = dims.num_nodes
    for (ss, ee) in qual[:20]:
        np.testing.assert_array_equal(f1[ss, 0, ee, :], f0[ss, 0, ee, :])
        np.testing.assert_allclose(f1[ss, n - 1, ee, :].sum(), f0[ss, n - 1, ee, :].sum(), rtol=1e-12)


def test_plain_check_reproduces_the_plain_pass():
    sol, _ = _solver(M=30)
    net = _network(30)
    try:
        sol.set_surrogate(16, 1e-3)
        f0 = sol.download_f()
        r = sol.surrogate_probe(1)
        f1 = sol.download_f()
    finally:
        net.close()
        sol.close()
    np.testing.assert_array_equal(f0, f1)  # the plain probe leaves f as it was
    print("largest sampled difference between the window solve and the plain pass: %.3e" % r)
    # the window solve takes the reference's operation order, the pass the
    # fused multiply-adds of precision mode 2
    assert r < 1e-12


def _run(sol, f0, steps, K=None, tol=None, graph=True):
    sol.upload_f(f0)
    sol.set_surrogate(0 if K is None else K, 1e-3 if tol is None else tol)
    sol.set_graph(graph)
    stats = []
    for _ in range(steps):
        sol.step()
        if K is not None:
            stats.append(sol.surrogate_stats())
    return sol.download_f(), stats


def test_zero_tolerance_is_the_plain_step():
    sol, dims = _solver(M=30)
    net = _network(30)
    try:
        f0 = sol.download_f()
        fp, _ = _run(sol, f0, 2)
        fs, st = _run(sol, f0, 2, K=8, tol=0.0)
    finally:
        net.close()
        sol.close()
    np.testing.assert_array_equal(fs, fp)
    for s in st:
        assert s["network_slabs"] == 0 and s["fallback_slabs"] > 0
        assert s["rejected_stream_substeps"] > 0


def test_no_tolerance_graph_equals_eager_and_counts_add_up():
    sol, dims = _solver(M=30)
    net = _network(30, scale=0.05)
    try:
        f0 = sol.download_f()
        fg, stg = _run(sol, f0, 2, K=8, tol=1e30, graph=True)
        fe, ste = _run(sol, f0, 2, K=8, tol=1e30, graph=False)
        fp, _ = _run(sol, f0, 2)
    finally:
        net.close()
        sol.close()
    np.testing.assert_array_equal(fg, fe)
    assert np.all(np.isfinite(fg)) and np.all(fg > 0)
    p = _params()
    for s in stg:
        assert s["fallback_slabs"] == 0 and s["rejected_stream_substeps"] == 0
        assert s["network_slabs"] > 0 and s["network_cell_subcycles"] > 0
        total = s["network_slabs"] + s["fallback_slabs"] + s["plain_slabs"]
        assert total == dims.num_streams * dims.num_energies * p.numEpSteps
    assert stg == ste
    # an untrained network moves f away from the plain result
    assert np.max(np.abs(fg - fp) / fp) > 1e-3


def test_surrogate_form_at_the_workload_size():
    """The surrogate form on a six-stream slice of the c5 workload at its
    per-stream size (2000 nodes, 50 energies, 30 mu, K = 16, the bench's
    nodes and parameters): with a zero tolerance every qualifying stream
    falls back and the step equals the plain step bit for bit; with the
    plain check off, every slab with 16 or more subcycles is advanced by
    the network and the counters add up."""
    c = workloads.WORKLOADS["c5"]
    p = workloads.workload_params(c)
    nodes = workloads.workload_nodes(c)
    sub = {k: np.ascontiguousarray(v[:6]) for k, v in nodes.items()}
    dims = workloads.dims_of(sub, c["E"], c["M"])
    sol = runtime.EpSolver(dims, p)
    net = _network(c["M"], scale=0.05)
    try:
        sol.set_nodes(sub)
        sol.init_seed_population()
        sol.set_precision(2)
        sol.step()
        f0 = sol.download_f()
        fp, _ = _run(sol, f0, 1)
        fs, st0 = _run(sol, f0, 1, K=c["surrogate"], tol=0.0)
        fn, st1 = _run(sol, f0, 1, K=c["surrogate"], tol=1e30)
    finally:
        net.close()
        sol.close()
    np.testing.assert_array_equal(fs, fp)
    assert st0[0]["network_slabs"] == 0 and st0[0]["fallback_slabs"] > 0
    s1 = st1[0]
    total = s1["network_slabs"] + s1["fallback_slabs"] + s1["plain_slabs"]
    assert total == dims.num_streams * dims.num_energies * p.numEpSteps
    assert s1["network_slabs"] + s1["fallback_slabs"] == st0[0]["fallback_slabs"]
    assert s1["network_slabs"] > 0
    assert np.all(np.isfinite(fn)) and np.all(fn > 0)

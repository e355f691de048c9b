"""Parity at the per-stream sizes of the benchmark workloads (through the C
ABI, against the FP64 oracle at 1e-10 relative on every cell of f):

* 384 x 2000 x 50 x 30: a slice of six streams in the arithmetic and with
  the tiling the benchmark line runs (precision mode 2, 768-node chunks,
  48-subcycle halo, three nodes per lane; N = 2000 makes three chunks per
  slab, so the cross-chunk halos are exercised), focusing in the last
  streaming pass or not;
* 1536 streams x 4000 x 64 x 32 with active-node counts drawn from
  200..4000: a six-stream slice with the workload's own ragged counts, and
  one global step of the whole workload (f is 2 x 101 GB on the device)
  whose last streams, at cell offsets far past 2^31, are compared with the
  oracle run on those streams alone (with the cross-stream operators off,
  streams are independent);
* 24 x 400 x 20 x 20 in full for one simulated day (64 global steps of 10
  EP substeps), with the per-step subcycle maxima;
* the two heavy-node focusing kernels (four lanes per row for a short
  heavy list, a lane per row for a long one) against each other, bit for
  bit, on the 384-stream nodes and a slice of them.
"""
import time

import numpy as np
import pytest

import oracle
from eprem_amd import runtime, workloads

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def log(msg):
    print("[workloads] " + msg, flush=True)


def subset(nodes, idx, n_active=None):
    sub = {k: np.ascontiguousarray(v[idx]) for k, v in nodes.items()}
    na = None if n_active is None else np.ascontiguousarray(n_active[idx])
    return sub, na


def max_rel(a, b, n_active=None):
    if n_active is not None:
        mask = np.zeros(a.shape[:2], bool)
        for s, n in enumerate(n_active):
            mask[s, :n] = True
        a, b = a[mask], b[mask]
    return float(np.max(np.abs(a - b) / np.abs(b)))


def device_step(nodes, n_active, E, M, p, tiling=None, precision=0, steps=1, f0=None):
    dims = workloads.dims_of(nodes, E, M)
    sol = runtime.EpSolver(dims, p)
    try:
        if precision:
            sol.set_precision(precision)
        if tiling:
            sol.set_tuning(*tiling)
        sol.set_nodes(nodes, n_active)
        if f0 is None:
            sol.init_seed_population()
            f0 = sol.download_f()
        else:
            sol.upload_f(f0)
        diags = []
        for _ in range(steps):
            _, dg = sol.step()
            diags.append(dg)
        return f0, sol.download_f(), diags
    finally:
        sol.close()


def oracle_steps(nodes, n_active, E, M, p, f0, steps=1):
    dims = workloads.dims_of(nodes, E, M)
    fo, diags = f0, []
    for _ in range(steps):
        rc, do, fo, _, _, _ = oracle.update_ep(dims, p, nodes, fo, n_active=n_active)
        assert rc == 0, rc
        diags.append(do)
    return fo, diags


def assert_diags(gd, od):
    for g, o in zip(gd, od):
        for k in ("max_subcycles_change", "max_subcycles_focusing", "max_nsteps_stream", "cell_updates"):
            assert g[k] == o[k], (k, g[k], o[k])
        assert g["min_tau"] == pytest.approx(o["min_tau"], rel=1e-14)


@pytest.mark.parametrize("fuse", [0, 1])
def test_c2_slice_at_the_benchmark_setting(fuse):
    c = workloads.WORKLOADS["c2"]
    p = workloads.workload_params(c)
    nodes = workloads.workload_nodes(c)
    idx = [0, 70, 141, 212, 290, 383]  # streams on every face
    sub, _ = subset(nodes, idx)
    log("c2 slice: 6 x 2000 x 50 x 30, mode 2, tiling 768/48/3 fuse %d" % fuse)
    f0, fg, gd = device_step(sub, None, c["E"], c["M"], p, tiling=(768, 48, 3, fuse), precision=2)
    assert gd[0]["max_nsteps_stream"] > 96  # three or more passes per slab
    t0 = time.time()
    fo, od = oracle_steps(sub, None, c["E"], c["M"], p, f0)
    log("oracle %.1f s, max rel %.3e" % (time.time() - t0, max_rel(fg, fo)))
    assert_diags(gd, od)
    assert max_rel(fg, fo) <= RTOL


@pytest.mark.parametrize("precision,tiling", [(2, (384, 32, 3, 0)), (0, (768, 48, 3, 1))])
def test_c3_slice_with_ragged_counts(precision, tiling):
    c = workloads.WORKLOADS["c3"]
    p = workloads.workload_params(c)
    S = 6 * c["rows"] * c["cols"]
    na_all = workloads.workload_active_counts(c, S)
    # the shortest and longest streams and four in between, from the
    # workload's own draw
    order = np.argsort(na_all)
    idx = [int(order[0]), int(order[S // 5]), int(order[S // 2]), int(order[4 * S // 5]), int(order[-2]),
           int(order[-1])]
    nodes = workloads.workload_nodes(c)
    sub, na = subset(nodes, idx, na_all)
    log("c3 slice: counts %s, mode %d, tiling %s" % (list(na), precision, tiling))
    f0, fg, gd = device_step(sub, na, c["E"], c["M"], p, tiling=tiling, precision=precision)
    fo, od = oracle_steps(sub, na, c["E"], c["M"], p, f0)
    log("max rel %.3e" % max_rel(fg, fo, na))
    assert_diags(gd, od)
    assert max_rel(fg, fo, na) <= RTOL


def test_c3_full_step_last_streams():
    c = workloads.WORKLOADS["c3"]
    p = workloads.workload_params(c)
    S = 6 * c["rows"] * c["cols"]
    na = workloads.workload_active_counts(c, S)
    log("c3 full: generating %d streams x %d nodes" % (S, c["N"]))
    nodes = workloads.workload_nodes(c)
    dims = workloads.dims_of(nodes, c["E"], c["M"])
    pick = [0, S - 3, S - 2, S - 1]
    cells_before = pick[1] * c["E"] * c["M"] * ((c["N"] + 63) // 64 * 64)
    assert cells_before > 2 ** 31
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_precision(2)
        sol.set_tuning(384, 32, 3, 0)
        sol.set_nodes(nodes, na)
        sol.init_seed_population()
        f0 = [sol.download_f_streams(s, 1)[0] for s in pick]
        log("c3 full: device step")
        _, dg = sol.step()
        log("c3 full: step done, %.1f ms" % dg["device_ms"])
        fg = [sol.download_f_streams(s, 1)[0] for s in pick]
        # the whole-range download agrees with the per-stream one
        tail = sol.download_f_streams(S - 3, 3)
    finally:
        sol.close()
    np.testing.assert_array_equal(tail, np.stack(fg[1:]))
    sub, na_sub = subset(nodes, pick, na)
    fo, _ = oracle_steps(sub, na_sub, c["E"], c["M"], p, np.stack(f0))
    err = max_rel(np.stack(fg), fo, na_sub)
    log("c3 full: streams %s (counts %s), max rel %.3e" % (pick, list(na_sub), err))
    assert err <= RTOL


@pytest.mark.parametrize("precision", [0, 2])
def test_c1_one_simulated_day(precision):
    c = workloads.WORKLOADS["c1"]
    p = workloads.workload_params(c)
    nodes = workloads.workload_nodes(c)
    steps = 64  # tDel = 1/64 day
    f0, fg, gd = device_step(nodes, None, c["E"], c["M"], p, precision=precision, steps=steps)
    t0 = time.time()
    fo, od = oracle_steps(nodes, None, c["E"], c["M"], p, f0, steps=steps)
    log("c1 day, mode %d: oracle %.1f s, max rel %.3e" % (precision, time.time() - t0, max_rel(fg, fo)))
    assert_diags(gd, od)
    assert max_rel(fg, fo) <= RTOL


@pytest.mark.parametrize("precision", [0, 2])
def test_heavy_focusing_kernels_give_the_same_bits(precision):
    """The heavy-node focusing walk runs four lanes per row for a short
    heavy list and a lane per row for a long one (k_focus_heavy4 /
    k_focus_heavy; the list length picks one on the device).  Streams are
    independent, so one step of the 384-stream workload at 8 energies x 16
    mu (a long list) and of a two-stream slice of it (a short one) must give
    the slice's streams bit for bit."""
    c = workloads.WORKLOADS["c2"]
    p = workloads.workload_params(c)
    nodes = workloads.workload_nodes(c)
    E, M = 8, 16
    idx = np.array([5, 200])
    out = []
    for nd in (nodes, subset(nodes, idx)[0]):
        dims = workloads.dims_of(nd, E, M)
        sol = runtime.EpSolver(dims, p)
        try:
            sol.set_precision(precision)
            sol.set_nodes(nd)
            sol.init_seed_population()
            _, dg = sol.step()
            heavy = dg["heavy_nodes"]
            if nd is nodes:
                out.append((heavy, np.concatenate([sol.download_f_streams(int(s), 1) for s in idx])))
            else:
                out.append((heavy, sol.download_f()))
        finally:
            sol.close()
    log("heavy nodes: %d (all streams), %d (slice)" % (out[0][0], out[1][0]))
    assert out[0][0] > 2048 and 0 < out[1][0] <= 2048
    np.testing.assert_array_equal(out[0][1], out[1][1])

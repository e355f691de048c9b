"""Benchmark of the EP update on one MI355X.

Workload (default, --config c2): 6x8x8 = 384 streams x 2000 nodes x 50
energies (10 keV - 1 GeV, uniform in ln p) x 30 mu, FP64, a Parker-spiral
snapshot (400 km/s, B_r = 2.5 nT at 1 AU, 25-day rotation) with seeded
perturbations, lambda = 0.1 AU (R/1 GV)^(1/3) (r/1 AU), seed spectrum
E^-2, tDel = 1/64 day, 10 EP substeps per global step.  A "step" is one
global step (ep_step: all 10 EP substeps of streaming + relaxation,
focusing and adiabatic change on every node).  The MAS trilinear
interpolation of a synthetic 300 x 200 x 400 FP32 grid (two time levels,
Parker solution + Gaussian CME shell) onto every node runs once per step
and is timed separately.

Prints one JSON line (metric, value, roofline, cpu_baseline, ...).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from eprem_amd import config, constants as K, mas_synth, runtime, workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector (FMA = 2 flops)
BF16_PEAK_TFLOPS = 2500.0    # MI355X dense bf16 MFMA
STREAM_FLOPS_PER_SWEEP = 8   # upwind 4 + iso sum 1 + relaxation 3 (FP64)

# HBM bytes per k_stream_pass launch, measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (each scaled by what the counter reads for a
# plain 1 GiB copy kernel: FETCH x 2.000, WRITE x 1.000), keyed by
# (workload, block, halo, nodes per lane, fused focusing, precision mode,
# schedule: 0 a launch per pass, 1 a persistent launch per substep); the commands and per-kernel
# figures are in profiles/k_stream_pass2_c2_counters.md (two nodes per lane)
# and profiles/k_stream_passR_c2_kernel_trace.md (three).  Other tilings
# report null.
STREAM_TRAFFIC = {("c2", 512, 48, 2, 0, 2, 0): 8.328e9, ("c2", 768, 48, 3, 0, 2, 0): 8.038e9}
# the same for one launch of the one-pass focusing + change kernel
# (profiles/k_focus_change4_c2_kernel_trace.md), keyed by (workload,
# precision mode)
NODE_TRAFFIC = {("c2", 2): 18.51e9}

CONFIGS = workloads.WORKLOADS


def run_surrogate(args, c, p, dims):
    """Learned-surrogate form of the EP update on the c2 workload.

    Training: >= 1e6 pairs (inputs: ln f at a node and its two neighbours
    + 6 scalars; target: ln f after K plain streaming subcycles and the
    focusing of the same time, as a correction to the large step that
    relaxes without transport) from four seeded wind/CME variants at two
    states each, trained on the device (forward, backward, Adam kernels).
    Timed: the same global steps from the same state, first on the plain
    path (FP64, mode 2, autotuned), then with ep_set_surrogate(K, 1e-3):
    network applications, the plain check that judges them, the per-stream
    fallback and the plain remainder are all inside the timed steps.
    Accuracy: omnidirectional flux at three observers after the steps,
    against the plain FP64 path (equal to the oracle to 1e-13, see the c2
    line's parity field)."""
    from eprem_amd import surrogate

    Ks = c["surrogate"]
    M = c["M"]
    n_active = active_counts(c, dims.num_streams)
    xs, ys = [], []
    variants = ((20261203, 0.5), (7, 0.3), (11, 0.8), (23, 0.6))
    for vi, (seed, cme_r) in enumerate(variants):
        nodes = bench_nodes(c, seed=seed, cme_r=cme_r)
        sol = runtime.EpSolver(dims, p)
        sol.set_nodes(nodes, n_active)
        sol.init_seed_population()
        sol.set_precision(2)
        sol.set_graph(False)
        for state in range(2):
            sol.step()  # anisotropy from the isotropic seed, then one step further
            x, y = sol.surrogate_samples(Ks, 260000, seed=seed * 10 + state)
            xs.append(x)
            ys.append(y)
            log("surrogate samples from variant %d state %d: %d" % (vi, state, len(x)))
        sol.close()
    x, y = np.concatenate(xs), np.concatenate(ys)
    del xs, ys
    rng = np.random.default_rng(1)
    order = rng.permutation(len(x))
    x, y = x[order], y[order]
    cut = len(x) - 50000
    t0 = time.perf_counter()
    net, hist = surrogate.train(x[:cut], surrogate.corrections(x[:cut], y[:cut], M), epochs=args.sur_epochs,
                                batch=8192, lr=2e-3)
    train_s = time.perf_counter() - t0
    log("trained on %d pairs: loss %.3e -> %.3e in %.1f s" % (cut, hist[0], hist[-1], train_s))
    pred = surrogate.predict(net, x[cut:], M)
    e_net = surrogate.flux_error(pred, y[cut:])
    e_base = surrogate.flux_error(surrogate.relaxation_step(x[cut:], M), y[cut:])
    npairs = int(len(x))
    del x, y, pred

    # the timed EP update: plain, then surrogate, from the same state
    nodes = bench_nodes(c)
    sol = runtime.EpSolver(dims, p)
    sol.set_nodes(nodes, n_active)
    sol.init_seed_population()
    sol.set_precision(2)
    blk, halo, r, fz, _ = sol.autotune()
    for _ in range(args.warmup):
        sol.step()
    f_start = sol.download_f()

    def timed(K=0, tol=1e-3):
        sol.upload_f(f_start)
        sol.set_surrogate(K, tol)
        sol.step()  # capture the step graph outside the timed region
        sol.upload_f(f_start)
        stats = []
        cells = 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _, dg = sol.step()
            cells += dg["cell_updates"]
            if K:
                stats.append(sol.surrogate_stats())
        wall = time.perf_counter() - t0
        _, flux = sol.observe(OBSERVERS["r_au"], OBSERVERS["theta"], OBSERVERS["phi"])
        return wall, cells, flux, stats

    wall_p, cells_p, flux_p, _ = timed()
    log("plain steps timed: %.1f ms per step" % (wall_p * 1e3 / args.steps))
    wall_s, cells_s, flux_s, st_s = timed(Ks, 1e-3)
    log("surrogate steps timed: %.1f ms per step" % (wall_s * 1e3 / args.steps))
    # the same without the check's fallback (a stream falls back only on a
    # non-finite network output): the network's own error, or where it
    # diverges
    try:
        wall_n, cells_n, flux_n, st_n = timed(Ks, 1e30)
        diverged = None
        log("surrogate steps without fallback timed: %.1f ms per step" % (wall_n * 1e3 / args.steps))
    except runtime.EpError as err:
        diverged = str(err)
        log("surrogate steps without fallback diverged: %s" % diverged)
    sol.close()
    net.close()

    def flux_err(fa, fb):
        rr = np.abs(fa - fb) / np.abs(fb)
        return {"max": float(rr.max()), "median": float(np.median(rr)), "within_1pct": float(np.mean(rr <= 0.01))}

    def tot(stats):
        keys = ("network_slabs", "fallback_slabs", "plain_slabs", "network_cell_subcycles",
                "rejected_stream_substeps")
        out = {k: int(sum(st[k] for st in stats)) for k in keys}
        out["max_residual"] = float(max(st["max_residual"] for st in stats))
        q = out["network_slabs"] + out["fallback_slabs"]
        out["fallback_fraction_of_network_slabs"] = out["fallback_slabs"] / q if q else None
        return out

    rate_p = cells_p / wall_p
    rate_s = cells_s / wall_s
    if diverged is None:
        rate_n = cells_n / wall_n
        nofb = {"value": rate_n, "ms_per_step": wall_n * 1e3 / args.steps, "over_plain": rate_n / rate_p,
                "stats": tot(st_n), "observer_flux_error_vs_fp64": flux_err(flux_n, flux_p)}
    else:
        nofb = {"diverged": diverged}
    sizes = [3 * M + 6, 128, 128, 128, M]
    flops_row = 2.0 * sum(a * b for a, b in zip(sizes[:-1], sizes[1:]))
    return {
        "metric": "phase-space cell-updates/s (1 GPU), surrogate vs plain at <=1% flux error",
        "value": rate_s, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall_s * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 MFMA network (fp32 accumulate), FP64 plain path",
        "data": "synthetic",
        "config": {"workload": "c5: c2 shapes (384 x 2000 x 50 x 30), MLP %d-128-128-128-%d replacing K=%d "
                               "streaming + focusing subcycles, residual tolerance 1e-3" % (3 * M + 6, M, Ks),
                   "path": "surrogate EP update with plain check and per-stream fallback (whole ep_step timed)",
                   "pairs": npairs, "train_s": train_s, "epochs": args.sur_epochs},
        "surrogate_over_plain": rate_s / rate_p,
        "plain": {"value": rate_p, "ms_per_step": wall_p * 1e3 / args.steps,
                  "tiling": {"block": blk, "halo": halo, "lane_nodes": r, "fuse_focus": fz}},
        "surrogate_stats": tot(st_s),
        "observer_flux_error_vs_fp64": flux_err(flux_s, flux_p),
        "without_fallback": nofb,
        "network_flop_per_cell_subcycle": flops_row / (M * Ks),
        "flux_error_heldout_pairs": {"network": surrogate.summary(e_net),
                                     "without_network_relaxation_only": surrogate.summary(e_base)},
        "loss_per_epoch": [float(v) for v in hist],
    }

OBSERVERS = dict(r_au=np.array([0.3, 0.5, 1.0]), theta=np.array([1.2, 1.571, 2.0]),
                 phi=np.array([0.5, 2.5, 4.5]))


def run_mixed(args, c, p, nodes, dims):
    """Mixed-precision form: the timed steps run with f stored as float
    (FP32 streaming, focusing and interpolation; FP64 relaxation and
    adiabatic change); the same steps are then repeated on the FP64 path to
    state the error of f, of the omnidirectional flux at three observers and
    of its fluence (time integral over the run), and on 16 streams against
    the FP64 oracle itself; then the same for the log-scaled variant (the
    float storage holding ln f)."""
    grids = None
    sol = runtime.EpSolver(dims, p)
    sol.set_nodes(nodes)
    grids = sol.grids()
    f0 = workloads.power_law_f(grids, nodes, c["M"], seed=20261203)
    tdel_days = 1.0 / 64

    def run_mode(sv, f_init, mode):
        sv.set_precision(0)
        sv.upload_f(f_init)
        sv.set_precision(mode)
        fluence = np.zeros((3, c["E"]))
        wall, cells = 0.0, 0
        flux = None
        for i in range(args.warmup + args.steps):
            t0 = time.perf_counter()
            _, dg = sv.step()
            t1 = time.perf_counter()
            if i >= args.warmup:
                wall += t1 - t0
                cells += dg["cell_updates"]
            _, flux = sv.observe(OBSERVERS["r_au"], OBSERVERS["theta"], OBSERVERS["phi"])
            fluence += flux * tdel_days
        return sv.download_f(), flux, fluence, wall, cells

    def run(mixed, timed):
        return run_mode(sol, f0, 1 if mixed else 0)

    tiling = None
    if not (args.block or args.halo):
        # tile search in the mode that is timed (float storage)
        sol.upload_f(f0)
        sol.set_precision(1)
        blk, halo, r, fz, _ = sol.autotune()
        tiling = {"block": blk, "halo": halo, "lane_nodes": r, "fuse_focus": fz, "schedule": sol.schedule()}
    f32, flux32, flu32, wall, cells = run(True, True)
    f64, flux64, flu64, wall64, cells64 = run(False, False)
    sol.close()
    # float storage floors at FLT_MIN: the error of f is stated over the
    # cells inside the workload's range (f >= 1e-30)
    live = f64 >= 1e-30
    rel = np.abs(f32[live] - f64[live]) / f64[live]
    live_frac = float(live.mean())
    f64k, live64 = f64, live
    del f32
    oracle_cmp = mixed_vs_oracle(args, c, p, nodes, grids, f0, tdel_days)
    log("mixed precision against the oracle done")
    # the log-scaled variant (precision mode 3: the float storage holds ln f)
    sol = runtime.EpSolver(dims, p)
    sol.set_nodes(nodes)
    if not (args.block or args.halo):
        sol.autotune()
    fl, fluxl, flul, wall_l, cells_l = run_mode(sol, f0, 3)
    sol.close()
    rel_l = np.abs(fl[live64] - f64k[live64]) / f64k[live64]
    del fl, f64k, live64
    log_cmp = mixed_vs_oracle(args, c, p, nodes, grids, f0, tdel_days, mode=3)
    log("log-scaled variant done")
    interp_cmp = None if args.no_mas else interp_fp32_vs_oracle(c, p)

    def err(a, b):
        r = np.abs(a - b) / np.abs(b)
        return {"max": float(r.max()), "median": float(np.median(r))}

    return {
        "metric": "phase-space cell-updates/s (1 GPU), surrogate vs plain at <=1% flux error",
        "value": cells / wall, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (change f64)", "data": "synthetic",
        "config": {"workload": "c4: c2 shapes, mixed precision, power-law f spanning %.1e..%.1e"
                               % (f0.min(), f0.max()), "path": "mixed precision"},
        "fp64_same_steps": {"value": cells64 / wall64, "ms_per_step": wall64 * 1e3 / args.steps},
        "tiling": tiling,
        "precision": "f stored as float; streaming (upwind), focusing and MAS interpolation in FP32; "
                     "relaxation (the pitch-angle part) and the adiabatic change in FP64",
        "error_vs_fp64_device_all_streams": {
            "f": {"max": float(rel.max()), "median": float(np.median(rel)), "cells_at_or_above_1e-30": live_frac},
            "observer_flux": err(flux32, flux64), "observer_fluence": err(flu32, flu64)},
        "error_vs_fp64_oracle": oracle_cmp,
        "interpolation_fp32_vs_oracle": interp_cmp,
        "log_scaled_variant": {
            "storage": "float holding ln f (precision mode 3); the operators as in the mixed form",
            "value": cells_l / wall_l, "ms_per_step": wall_l * 1e3 / args.steps,
            "error_vs_fp64_device_all_streams": {"f": {"max": float(rel_l.max()),
                                                       "median": float(np.median(rel_l))},
                                                 "observer_flux": err(fluxl, flux64),
                                                 "observer_fluence": err(flul, flu64)},
            "error_vs_fp64_oracle": log_cmp},
    }


def mixed_vs_oracle(args, c, p, nodes, grids, f0, tdel_days, nsub=16, mode=1):
    """The mixed-precision form against the FP64 oracle itself on the first
    nsub streams of the workload, over the same steps: relative error of f
    (cells >= 1e-30), of the omnidirectional flux at the three observers
    (both blended over the same nsub streams) and of its fluence."""
    import oracle

    cores, _ = host_cores()
    sub = {k: np.ascontiguousarray(v[:nsub]) for k, v in nodes.items()}
    sdims = workloads.dims_of(sub, c["E"], c["M"])
    f0s = np.ascontiguousarray(f0[:nsub])
    nsteps = args.warmup + args.steps
    obs = (OBSERVERS["r_au"], OBSERVERS["theta"], OBSERVERS["phi"])
    fo = f0s.copy()
    flu_o = np.zeros((3, c["E"]))
    flux_o = None
    for _ in range(nsteps):
        rc, _, fo, _, _, _ = oracle.update_ep(sdims, p, sub, fo, nthreads=cores)
        if rc != 0:
            raise RuntimeError("oracle failed: %d" % rc)
        _, flux_o = oracle.observers(sdims, sub, fo, grids["egrid"], p.rScale, *obs)
        flu_o += flux_o * tdel_days
    ss = runtime.EpSolver(sdims, p)
    try:
        ss.set_nodes(sub)
        ss.upload_f(f0s)
        ss.set_precision(mode)
        flu_m = np.zeros((3, c["E"]))
        flux_m = None
        for _ in range(nsteps):
            ss.step()
            _, flux_m = ss.observe(*obs)
            flu_m += flux_m * tdel_days
        fm = ss.download_f()
    finally:
        ss.close()
    live = fo >= 1e-30
    rel = np.abs(fm[live] - fo[live]) / fo[live]

    def err(a, b):
        r = np.abs(a - b) / np.abs(b)
        return {"max": float(r.max()), "median": float(np.median(r))}

    return {"streams": nsub, "global_steps": nsteps,
            "f": {"max": float(rel.max()), "median": float(np.median(rel)),
                  "cells_at_or_above_1e-30": float(live.mean())},
            "observer_flux": err(flux_m, flux_o), "observer_fluence": err(flu_m, flu_o)}


def interp_fp32_vs_oracle(c, p, nsub=16):
    """MAS trilinear interpolation in FP32 (the mixed-precision form) and in
    FP64, each against the oracle's masTriLinear, on radial streams from 1.2
    to 29 Rs (all inside the synthetic MAS domain): relative error of each
    of the seven outputs, scaled by that output's largest magnitude."""
    import oracle

    lv0, lv1 = mas_synth.levels(seed=20261203)
    S = 6 * c["rows"] * c["cols"]
    N = c["N"]
    i = np.arange(S) + 0.5
    ct = 1.0 - 2.0 * i / S
    ph = np.pi * (1.0 + 5.0 ** 0.5) * i
    rs = K.RSAU / p.rScale
    r = np.linspace(1.2, 29.0, N)[None, :] * rs * np.ones((S, 1))
    st = np.sqrt(1.0 - ct ** 2)
    x, y, z = r * (st * np.cos(ph))[:, None], r * (st * np.sin(ph))[:, None], r * ct[:, None]
    nodes = dict(bench_nodes(c), x=x, y=y, z=z, rmag=r)
    dims = workloads.dims_of(nodes, 2, 2)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_nodes(nodes)
        sol.mas_set_level(0, *lv0)
        sol.mas_set_level(1, *lv1)
        prev = mas_synth.prev_from_nodes(nodes, p.rScale)
        out = {}
        for mode in (0, 1):
            sol.set_precision(mode)
            out[mode] = sol.interp_mhd(0.5, 0.0, 30.0, 50.0e5, prev)
    finally:
        sol.close()
    sub = slice(0, nsub)
    ref = oracle.mas_interp(p.rScale, x[sub].ravel(), y[sub].ravel(), z[sub].ravel(), r[sub].ravel(), lv0, lv1,
                            0.5, 0.0, 30.0, 50.0e5)
    res = {}
    for name, mode in (("fp32", 1), ("fp64", 0)):
        comp = {}
        for k in ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density"):
            a = np.asarray(out[mode][k])[sub].ravel()
            b = ref[k]
            scale = np.abs(b).max()
            comp[k] = float(np.abs(a - b).max() / scale) if scale > 0 else float(np.abs(a).max())
        res[name] = comp
    res["nodes_compared"] = int(nsub * N)
    res["note"] = "max |device - oracle| / max |oracle| per output (max |device| where the oracle's output is 0)"
    return res


def log(msg):
    """Progress on stderr (the JSON result goes to stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def active_counts(c, S):
    return workloads.workload_active_counts(c, S)


def make_params(c):
    return workloads.workload_params(c)


def bench_nodes(c, seed=20261203, cme_r=0.5):
    return workloads.workload_nodes(c, seed=seed, cme_r=cme_r)


MAS_BYTES_PER_NODE = 504.0  # 7 fields x 2 time levels x 8 FP32 corners read + 7 FP64 written


def interp_inside(c, p, lv0, lv1, reps):
    """masTriLinear on every node: the same stream count and length as the
    workload, but each stream a radial line from 1.2 to 29 Rs (inside the
    synthetic MAS domain, 1-30 Rs) in its own direction (a Fibonacci
    sphere), so that no node takes the masWind branch.  Returns the kernel
    time and its rate against the 504 algorithmic bytes per node."""
    S = 6 * c["rows"] * c["cols"]
    N = c["N"]
    rng = np.random.default_rng(7)
    i = np.arange(S) + 0.5
    ct = 1.0 - 2.0 * i / S
    ph = np.pi * (1.0 + 5.0 ** 0.5) * i
    rs = K.RSAU / p.rScale  # one solar radius in code length units
    r = np.linspace(1.2, 29.0, N)[None, :] * rs * (1.0 + 0.002 * rng.standard_normal((S, N)))
    st = np.sqrt(1.0 - ct ** 2)
    x, y, z = r * (st * np.cos(ph))[:, None], r * (st * np.sin(ph))[:, None], r * ct[:, None]
    nodes = bench_nodes(c)
    nodes = dict(nodes, x=x, y=y, z=z, rmag=r)
    dims = workloads.dims_of(nodes, 2, 2)
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_nodes(nodes)
        sol.mas_set_level(0, *lv0)
        sol.mas_set_level(1, *lv1)
        prev = mas_synth.prev_from_nodes(nodes, p.rScale)
        out = sol.interp_mhd(0.5, 0.0, 30.0, 50.0e5, prev)
        ms = min(sol.interp_mhd(0.5, 0.0, 30.0, 50.0e5, prev)["device_ms"] for _ in range(reps))
    finally:
        sol.close()
    assert np.isfinite(out["br"]).all()
    nn = S * N
    return {"kernel": "k_mas_interp", "nodes": nn, "all_inside_mas_domain": True, "ms": ms,
            "achieved_gbs": nn * MAS_BYTES_PER_NODE / (ms * 1e-3) / 1e9, "peak_gbs": HBM_PEAK_GBS,
            "frac": nn * MAS_BYTES_PER_NODE / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "bytes_per_node": MAS_BYTES_PER_NODE}


def host_cores():
    """Threads the CPU baseline may use: the cores this process may run on
    (its affinity set, the box's share), capped by OMP_NUM_THREADS when set;
    with what the host reports beside it."""
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    n = aff
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return n, {"nproc": os.cpu_count(), "affinity_cores": aff, "omp_num_threads": omp, "cpu_model": model}


def cpu_baseline(c, p, nodes, n_active=None, budget_s=10.0):
    """The FP64 oracle (OpenMP over streams) on a bounded stream subset of
    the same workload, scaled per cell-update: one global step on the
    subset (at least one stream per thread), repeated from the same state
    until about budget_s seconds of CPU work have been timed.  Returns the
    baseline record and, for the parity check, the subset, its initial f
    and the oracle's f after the step."""
    import oracle

    cores, host = host_cores()
    E, M = c["E"], c["M"]
    nsub = max(16, cores)
    sub = {k: np.ascontiguousarray(v[:nsub]) for k, v in nodes.items()}
    na = None if n_active is None else np.ascontiguousarray(n_active[:nsub])
    dims = workloads.dims_of(sub, E, M)
    g = oracle.build_grids(p, E, M)
    f0 = workloads.seeded_f(p, g, sub, noise=0.0)
    dt, cells, reps = 0.0, 0, 0
    f1 = None
    while reps == 0 or (dt < budget_s and reps < 200):
        t0 = time.perf_counter()
        rc, d, fo, _, _, _ = oracle.update_ep(dims, p, sub, f0, n_active=na, nthreads=cores)
        dt += time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError("oracle failed: %d" % rc)
        if f1 is None:
            f1 = fo
        cells += d["cell_updates"]
        reps += 1
    rec = {"value": cells / dt, "unit": "cell-updates/s", "cores": cores, "kind": "port",
           "host": host,
           "sample": "%d of %d streams x %d nodes x %d E x %d mu, 1 global step (%d EP substeps) "
                     "repeated %d times from the same state, %.1f s, %d OpenMP threads" % (
                         nsub, 6 * c["rows"] * c["cols"], c["N"], E, M, p.numEpSteps, reps, dt, cores)}
    return rec, (sub, na, f0, f1)


def parity_check(c, p, tiling, precision, sample):
    """Max relative error of f against the oracle after one global step on
    the CPU baseline's streams, run on the device with the benchmark's
    precision mode and tiling (outside the timed region)."""
    sub, na, f0, f1 = sample
    dims = workloads.dims_of(sub, c["E"], c["M"])
    sol = runtime.EpSolver(dims, p)
    try:
        sol.set_precision(precision)
        sol.set_tuning(tiling["block"], tiling["halo"], tiling["lane_nodes"], tiling["fuse_focus"])
        sol.set_nodes(sub, na)
        sol.upload_f(f0)
        sol.step()
        fg = sol.download_f()
    finally:
        sol.close()
    mask = np.ones(fg.shape[:2], bool)
    if na is not None:
        for s, n in enumerate(na):
            mask[s, n:] = False
    rel = np.abs(fg[mask] - f1[mask]) / np.abs(f1[mask])
    return {"max_rel_err_f": float(rel.max()), "median_rel_err_f": float(np.median(rel)),
            "tolerance": 1e-10, "streams": int(dims.num_streams),
            "against": "FP64 oracle, one global step from the seed population"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--halo", type=int, default=0)
    ap.add_argument("--lane-nodes", type=int, default=1, choices=(1, 2, 3))
    ap.add_argument("--fuse-focus", type=int, default=1, choices=(0, 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mas", action="store_true")
    ap.add_argument("--no-unbatched", action="store_true")
    ap.add_argument("--sur-epochs", type=int, default=12, help="c5: training epochs")
    ap.add_argument("--schedule", type=int, default=-1, choices=(-1, 0, 1, 2),
                    help="with --block/--halo: streaming schedule (0 a launch per pass, 1 persistent, "
                         "2 per stream); -1 keeps the library's default")
    ap.add_argument("--precision", type=int, default=2, choices=(0, 2),
                    help="0: FP64 in the reference's operation order; 2: FP64 with fused "
                         "multiply-adds in streaming (within 1e-13 of it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the library runs every stream on one GPU; --gpus %d is not supported "
              "(the path would shard by stream: each GPU would own a block of streams and "
              "their f)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    c = CONFIGS[args.config]
    p = make_params(c)
    nodes = bench_nodes(c)
    dims = workloads.dims_of(nodes, c["E"], c["M"])
    if c.get("mixed"):
        print(json.dumps(run_mixed(args, c, p, nodes, dims)))
        return
    if c.get("surrogate"):
        print(json.dumps(run_surrogate(args, c, p, dims)))
        return
    n_active = active_counts(c, dims.num_streams)
    sol = runtime.EpSolver(dims, p)
    sol.set_nodes(nodes, n_active)
    sol.init_seed_population()
    sol.set_precision(args.precision)
    log("nodes and seed population set; precision mode %d" % args.precision)
    if args.block or args.halo:
        sol.set_tuning(args.block or 256, args.halo or 32, args.lane_nodes, args.fuse_focus)
        if args.schedule >= 0:
            sol.set_schedule(args.schedule)
        tiling = {"block": args.block or 256, "halo": args.halo or 32, "lane_nodes": args.lane_nodes,
                  "fuse_focus": args.fuse_focus, "source": "flags"}
    else:
        blk, halo, r, fz, tune_ms = sol.autotune()
        log("autotune: block %d halo %d lane_nodes %d fuse_focus %d" % (blk, halo, r, fz))
        tiling = {"block": blk, "halo": halo, "lane_nodes": r, "fuse_focus": fz, "source": "autotune",
                  "stream_and_focus_ms_per_substep": tune_ms}
    tiling["schedule"] = {0: "one launch per pass", 1: "one persistent launch per substep",
                          2: "one launch per stream"}[sol.schedule()]
    sched = sol.schedule()
    # the timed steps replay the captured step graph (the library's default);
    # the same steps with the kernels launched one by one, which lets HIP
    # events bracket each operator, give the per-operator breakdown

    mas_ms = None
    if not args.no_mas:
        lv0, lv1 = mas_synth.levels(seed=20261203)
        sol.mas_set_level(0, *lv0)
        sol.mas_set_level(1, *lv1)
        prev = mas_synth.prev_from_nodes(nodes, p.rScale)

    for _ in range(args.warmup):
        sol.step()
        log("warmup step done")
        if not args.no_mas:
            sol.interp_mhd(0.5, 0.0, 30.0, 50.0e5, prev)

    acc = dict(device_ms=0.0, stream_ms=0.0, focus_ms=0.0, change_ms=0.0, prep_ms=0.0,
               stream_sweeps=0, focus_sweeps=0, change_sweeps=0, stream_launches=0, cell_updates=0)
    mas_total = 0.0
    cells = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, dg = sol.step()
        cells += dg["cell_updates"]
    t1 = time.perf_counter()
    log("timed steps done (graph)")
    sol.set_graph(False)
    te0 = time.perf_counter()
    for _ in range(args.steps):
        _, dg = sol.step()
        for k in acc:
            acc[k] += dg[k]
    te1 = time.perf_counter()
    sol.set_graph(True)
    unbatched = None
    if c.get("ragged") and sched == 1 and not args.no_unbatched:
        # the unbatched form of the same persistent kernel: one launch per
        # stream per substep, the reference's loop over streams
        sol.set_schedule(2)
        sol.set_graph(False)
        sol.step()
        tu0 = time.perf_counter()
        _, du = sol.step()
        tu1 = time.perf_counter()
        unbatched = {"ms_per_step": (tu1 - tu0) * 1e3, "stream_ms_per_step": du["stream_ms"],
                     "stream_launches_per_step": du["stream_launches"],
                     "batched_stream_ms_per_step": acc["stream_ms"] / args.steps,
                     "batched_over_unbatched_speedup": (tu1 - tu0) * 1e3 / ((t1 - t0) * 1e3 / args.steps)}
        sol.set_schedule(1)
        sol.set_graph(True)
        log("unbatched comparison done")
    if not args.no_mas:
        for _ in range(args.steps):
            mas_total += sol.interp_mhd(0.5, 0.0, 30.0, 50.0e5, prev)["device_ms"]
        mas_ms = mas_total / args.steps
    wall = t1 - t0
    value = cells / wall

    # dominant kernel: the streaming passes
    kname = {1: "k_stream_pass", 2: "k_stream_pass2", 3: "k_stream_passR<R=3>"}[tiling["lane_nodes"]]
    if sched:
        kname = "k_stream_slab<R=%d>" % tiling["lane_nodes"]
    st_s = acc["stream_ms"] / 1e3
    launches = max(1, acc["stream_launches"])
    alg_bytes_per_launch = 16.0 * acc["stream_sweeps"] / launches
    avg_launch_s = st_s / launches
    achieved = alg_bytes_per_launch / avg_launch_s / 1e9
    fp64 = STREAM_FLOPS_PER_SWEEP * acc["stream_sweeps"] / st_s / 1e12
    traffic = STREAM_TRAFFIC.get((args.config, tiling["block"], tiling["halo"], tiling["lane_nodes"],
                                  tiling["fuse_focus"], args.precision, sched))
    # node stage (focusing + adiabatic change of each node in one pass):
    # HBM bound, one read and one write of every active cell per substep
    node_roof = None
    if dg["max_subcycles_change"] <= 1 and acc["change_ms"] > 0:
        node_bytes = 16.0 * acc["cell_updates"]
        node_gbs = node_bytes / (acc["change_ms"] * 1e-3) / 1e9
        node_roof = {"bound": "hbm", "kernel": "k_focus_change4" if c["M"] >= 12 else "k_focus_change",
                     "achieved": node_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": node_gbs / HBM_PEAK_GBS,
                     "traffic": NODE_TRAFFIC.get((args.config, args.precision)),
                     "algorithmic_bytes_per_launch": node_bytes / (args.steps * p.numEpSteps)}
    res = {
        "metric": "phase-space cell-updates/s (1 GPU), surrogate vs plain at <=1% flux error",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "%s: %d streams x %d nodes x %d energies x %d mu, FP64 plain path, "
                               "%d EP substeps per step" % (args.config, dims.num_streams, c["N"], c["E"],
                                                             c["M"], c["eps"]),
                   "path": "plain FP64 (no surrogate)",
                   "arithmetic": ("FP64, reference operation order" if args.precision == 0 else
                                  "FP64, fused multiply-add streaming (within 1e-13 of the reference order)")},
        "roofline": {"bound": "fp64", "kernel": kname, "achieved": fp64, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": fp64 / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "note": "FP64 vector issue binds the streaming kernel: %d algorithmic flops per "
                             "cell-sweep (upwind 4, isotropic sum 1, relaxation 3); it keeps up to `halo` "
                             "sweeps on chip, so its HBM traffic is far below the 16 B per cell-sweep of "
                             "one read and one write per sweep" % STREAM_FLOPS_PER_SWEEP,
                     "algorithmic_bytes_per_launch": alg_bytes_per_launch,
                     "avg_launch_ms": avg_launch_s * 1e3,
                     "hbm_gbs_measured_traffic": (traffic / avg_launch_s / 1e9) if traffic else None},
        "node_stage_roofline": node_roof,
        "breakdown_ms_per_step": {k: acc[k] / args.steps for k in
                                  ("device_ms", "prep_ms", "stream_ms", "focus_ms", "change_ms")},
        "node_stage": ("focus_ms: heavy-node focusing (k_focus_heavy4 for up to 2048 heavy nodes, else k_focus_heavy); change_ms: focusing and adiabatic "
                       "change in one pass (k_focus_change4 from 12 mu cells, k_focus_change below)") if dg["max_subcycles_change"] <= 1 else
                      "focus_ms: k_focus + k_focus_heavy; change_ms: k_change",
        "sweeps_per_step": {k: acc[k] // args.steps for k in ("stream_sweeps", "focus_sweeps",
                                                              "change_sweeps")},
        "interp_ms": mas_ms,
        "interp_nodes_inside_mas_domain": float(np.mean(nodes["rmag"] * p.rScale / K.RSAU < 30.0)),
        "tiling": tiling,
        "max_subcycles": {k: dg[k] for k in ("max_nsteps_stream", "max_subcycles_focusing",
                                             "max_subcycles_change")},
        "eager_ms_per_step": (te1 - te0) * 1e3 / args.steps,
        "eager_note": "the same steps with kernels launched one by one and HIP events around each operator "
                      "(the breakdown and roofline figures come from these)",
    }
    if unbatched is not None:
        res["unbatched"] = unbatched
    if n_active is not None:
        res["config"]["active_nodes"] = {"min": int(n_active.min()), "max": int(n_active.max()),
                                         "mean": float(n_active.mean())}
    if not args.no_mas:
        sol.close()
        sol = None
        log("MAS interpolation with every node inside the MAS domain")
        res["interp_inside"] = interp_inside(c, p, lv0, lv1, args.steps)
    if not args.no_cpu:
        log("timed steps done; CPU baseline")
        res["cpu_baseline"], sample = cpu_baseline(c, p, nodes, n_active)
        log("parity of the benchmark setting on the baseline's streams")
        res["parity"] = parity_check(c, p, tiling, args.precision, sample)
    if sol is not None:
        sol.close()
    print(json.dumps(res))


if __name__ == "__main__":
    main()

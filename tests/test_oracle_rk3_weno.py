"""Pins of the oracle's RK3 and WENO3 operator variants (adiabaticFocusAlg /
adiabaticChangeAlg = 2, 3; src/energeticParticles.c:952-1018, :1179-1427,
:1677-1743, :1899-2144) by what the mathematics fixes:

* the three-stage Shu-Osher step applied to a linear operator L is exactly
  the cubic Taylor polynomial I + dt L + (dt L)^2/2 + (dt L)^3/6, which pins
  the stage weights (3/4, 1/4; 1/3, 2/3) through a whole focusing step of
  the upwind operator (alg 2), with L probed from the operator itself;
* the reference's WENO3 flux (split fluxes LP = v f / 2 - alpha, LN =
  v f / 2 + alpha, two-point reconstructions, nonlinear weights) is exact
  for a linear profile under a constant velocity, which pins its stencils;
* at a face where one of the two candidate stencils is flat and the other
  is not, the nonlinear weights select the flat one completely;
* the focusing operator conserves sum_mu f (zero-flux walls), with WENO3
  too, and so does a whole RK3 + WENO3 focusing step;
* zero velocity leaves f unchanged;
* under a constant velocity at the subcycle's Courant number (0.9) the RK3
  upwind subcycle makes no new extrema and does not raise the total
  variation of a compact profile (a downwind flux does both).
The reference's WENO3 splits off alpha rather than alpha f, so it is not a
third-order scheme (the error of d(v f)/dmu on a smooth profile falls only
about linearly with the spacing); no convergence order is claimed for it.
"""
import math

import numpy as np
import pytest

import oracle
from eprem_amd import abi, config, workloads


def params(**kw):
    base = dict(numEpSteps=2, tDel=1.0 / 64, eMin=0.01, eMax=1000.0, lamo=0.1, mfpRadialPower=1.0,
                boundaryFunctEcutoff=1.0e30)
    base.update(kw)
    return config.default_params(**base)


def probe_matrix(op, n_rows, n_cols):
    """Matrix of a linear operator on [n_rows][n_cols] arrays."""
    n = n_rows * n_cols
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = op(e.reshape(n_rows, n_cols)).ravel()
    return A


def focus_v_avg(p, g, ds, bplus, bminus, dlnn, dlnb, dupar):
    """Half-mesh focusing velocities of a node (AdiabaticFocusing,
    src/energeticParticles.c:1581-1653), the same for every operator."""
    dt_full = p.tDel
    dupardt = dupar / dt_full
    b = 2.0 * (dlnn / dt_full) - 3.0 * (dlnb / dt_full)
    dlnbds = 0.0 if (bplus == 0.0 or bminus == 0.0) else (math.log(bplus) - math.log(bminus)) / (2.0 * ds + 2.2250738585072014e-308)
    E, M = len(g["vgrid"]), len(g["mugrid"])
    va = np.zeros((E, M + 1))
    for e in range(E):
        v = g["vgrid"][e]
        a = -v * dlnbds - (2.0 / v) * dupardt
        vel = 0.5 * (1.0 - g["mugrid"] ** 2) * (a + g["mugrid"] * b)
        va[e, 1:M] = 0.5 * (vel[1:] + vel[:-1])
        va[e, 0], va[e, M] = va[e, 1], va[e, M - 1]
    return va


def test_rk3_focusing_step_is_the_cubic_taylor_polynomial():
    E, M = 4, 9
    p1, p2 = params(adiabaticFocusAlg=1), params(adiabaticFocusAlg=2)
    nodes = workloads.parker_nodes(1, 1, 12, seed=5, perturb=0.3, jitter=0.2)
    dims = workloads.dims_of(nodes, E, M)
    g = oracle.build_grids(p1, E, M)
    f0 = workloads.seeded_f(p1, g, nodes, noise=0.4, seed=6)
    rc, d1, _, ds, bp, bm = oracle.update_ep(dims, p1, nodes, f0, op_flags=abi.EP_OP_FOCUS)
    assert rc == 0
    rc, d2, f2, _, _, _ = oracle.update_ep(dims, p2, nodes, f0, op_flags=abi.EP_OP_FOCUS)
    assert rc == 0
    dt = p2.tDel / p2.numEpSteps
    s, k = 2, 5
    va = focus_v_avg(p2, g, ds[s, k], bp[s, k], bm[s, k], nodes["dlnn"][s, k], nodes["dlnb"][s, k],
                     nodes["dupar"][s, k])
    A = -probe_matrix(lambda x: oracle.focus_operator(1, x, va, g["dmu"]), E, M)
    n_sub = int(math.ceil(dt / (0.9 * g["dmu"] / np.abs(va[:, 1:M]).max())))
    h = dt / n_sub
    P = np.eye(E * M) + h * A + (h * A) @ (h * A) / 2 + (h * A) @ (h * A) @ (h * A) / 6
    f = f0[s, k].ravel()
    for _ in range(p2.numEpSteps):  # the global step's EP substeps, each n_sub RK3 subcycles
        for _ in range(n_sub):
            f = P @ f
        f = np.maximum(f, 2.2250738585072014e-308)
    np.testing.assert_allclose(f2[s, k].ravel(), f, rtol=1e-12)
    # and the forward-Euler subcycles of alg 1 are (I + h A)^n
    rc, _, f1, _, _, _ = oracle.update_ep(dims, p1, nodes, f0, op_flags=abi.EP_OP_FOCUS)
    fe = f0[s, k].ravel()
    for _ in range(p1.numEpSteps):
        for _ in range(n_sub):
            fe = fe + h * (A @ fe)
        fe = np.maximum(fe, 2.2250738585072014e-308)
    np.testing.assert_allclose(f1[s, k].ravel(), fe, rtol=1e-12)
    assert np.max(np.abs(f - fe) / fe) > 1e-9  # the two schemes differ


@pytest.mark.parametrize("v", [0.37, -0.21])
def test_weno3_focusing_exact_for_linear_profiles(v):
    M, dmu = 20, 0.1
    mu = dmu * (np.arange(M) + 0.5)
    f = np.stack([2.0 + 0.7 * mu, 5.0 - 1.3 * mu])
    va = np.full((2, M + 1), v)
    f1 = oracle.focus_operator(3, f, va, dmu)
    # faces 2..M-2 are WENO3 faces; cells 2..M-3 use two of them
    np.testing.assert_allclose(f1[0, 2:M - 2], v * 0.7, rtol=1e-12)
    np.testing.assert_allclose(f1[1, 2:M - 2], v * -1.3, rtol=1e-12)
    # cells 1 and M-2 pair a WENO3 face (v f at the face: the alpha terms
    # cancel) with the upwind face next to the wall (v f of the upwind cell)
    for row, slope in ((0, 0.7), (1, -1.3)):
        fr = f[row]
        up1 = fr[0] if v > 0 else fr[1]
        upM = fr[M - 2] if v > 0 else fr[M - 1]
        face2, faceM2 = 0.5 * (fr[1] + fr[2]), 0.5 * (fr[M - 3] + fr[M - 2])
        assert f1[row, 1] == pytest.approx(v * (face2 - up1) / dmu, rel=1e-12)
        assert f1[row, M - 2] == pytest.approx(v * (upM - faceM2) / dmu, rel=1e-12)


@pytest.mark.parametrize("alg", [1, 3])
def test_focusing_operators_are_mirror_symmetric(alg):
    """mu -> -mu: reversing the cells and faces and negating the velocity
    maps d(v f)/dmu onto itself at the mirrored cell."""
    rng = np.random.default_rng(11)
    E, M, dmu = 2, 13, 2.0 / 13
    f = rng.uniform(0.5, 2.0, (E, M))
    va = rng.uniform(-1.0, 1.0, (E, M + 1))
    va[:, 0], va[:, M] = va[:, 1], va[:, M - 1]
    f1 = oracle.focus_operator(alg, f, va, dmu)
    g1 = oracle.focus_operator(alg, f[:, ::-1].copy(), -va[:, ::-1].copy(), dmu)
    np.testing.assert_allclose(g1[:, ::-1], f1, rtol=1e-12, atol=1e-14 * np.abs(f1).max())


def test_weno3_focusing_selects_the_flat_stencil():
    """A profile symmetric about face k (f[k-1] = f[k], slope -1 on the
    left, +1 on the right).  Of the two candidate stencils of each split
    flux, the one holding cells k-1 and k is flat (smoothness indicator 0)
    and takes all the weight, so with constant v (alpha cancels):
    F_k = v f[k]; F_{k-1} = v (f[k-1] + f(face k-1)) / 2 (LN from the
    straight left side is exact, LP takes the flat stencil); F_{k+1} the
    mirror image."""
    M, dmu, v, k = 16, 0.125, 0.4, 8
    x = dmu * (np.arange(M) + 0.5)
    f = (1.0 + np.abs(x - k * dmu))[None, :]
    va = np.full((1, M + 1), v)
    f1 = oracle.focus_operator(3, f, va, dmu)[0]
    fk = f[0, k]
    face = 1.0 + dmu  # f at faces k-1 and k+1
    F_k, F_side = v * fk, 0.5 * v * (fk + face)
    np.testing.assert_allclose(f1[k - 1], (F_k - F_side) / dmu, rtol=1e-12)
    np.testing.assert_allclose(f1[k], (F_side - F_k) / dmu, rtol=1e-12)
    # with equal weights instead of all-or-nothing selection both values
    # would differ (the indicator of the one-sided stencils is 4 (v/4)^2)
    assert abs((F_k - F_side) / dmu - (v * fk - v * face) / dmu) > 0.05


@pytest.mark.parametrize("alg", [1, 3])
def test_focusing_operators_conserve_sum_over_mu(alg):
    rng = np.random.default_rng(3)
    E, M, dmu = 3, 17, 2.0 / 17
    f = rng.uniform(0.5, 2.0, (E, M))
    va = rng.uniform(-1.0, 1.0, (E, M + 1))
    va[:, 0], va[:, M] = va[:, 1], va[:, M - 1]
    f1 = oracle.focus_operator(alg, f, va, dmu)
    scale = np.abs(va[:, 1:M] * f[:, 1:]).sum(axis=1) / dmu
    assert np.all(np.abs(f1.sum(axis=1)) <= 1e-14 * scale)
    zero = oracle.focus_operator(alg, f, np.zeros_like(va), dmu)
    assert np.all(zero == 0.0)


@pytest.mark.parametrize("alg", [2, 3])
def test_rk3_focusing_step_conserves_and_differs(alg):
    E, M = 6, 12
    p = params(adiabaticFocusAlg=alg, numEpSteps=1)
    p1 = params(adiabaticFocusAlg=1, numEpSteps=1)
    nodes = workloads.parker_nodes(1, 1, 16, seed=7, perturb=0.3, jitter=0.2)
    dims = workloads.dims_of(nodes, E, M)
    g = oracle.build_grids(p, E, M)
    # values of 1e6: the reference's WENO3 split fluxes carry alpha (a
    # velocity) unweighted by f, which on small values of f drives cells
    # negative; they are then floored at DBL_MIN, which adds to sum_mu f
    rng = np.random.default_rng(8)
    f0 = 1.0e6 * np.exp(rng.normal(0, 0.3, (dims.num_streams, dims.num_nodes, E, M)))
    rc, d, f, _, _, _ = oracle.update_ep(dims, p, nodes, f0, op_flags=abi.EP_OP_FOCUS)
    assert rc == 0 and d["max_subcycles_focusing"] >= 1
    _, _, fu, _, _, _ = oracle.update_ep(dims, p1, nodes, f0, op_flags=abi.EP_OP_FOCUS)
    # rows where the floor never acted keep their sum exactly (strong
    # focusing drives most WENO3 rows negative somewhere, and the floor
    # then adds to the sum)
    kept = f.min(axis=3) > 2.2250738585072014e-308
    assert kept[:, 1:].mean() > (0.9 if alg == 2 else 0.1)
    np.testing.assert_allclose(f.sum(axis=3)[:, 1:][kept[:, 1:]], f0.sum(axis=3)[:, 1:][kept[:, 1:]], rtol=1e-12)
    np.testing.assert_array_equal(f[:, 0], f0[:, 0])  # node 0 is not a compute shell
    assert np.max(np.abs(f - fu) / fu) > 1e-8


@pytest.mark.parametrize("v", [0.8, -0.6])
def test_weno3_change_exact_for_linear_profiles(v):
    """ln f linear in ln p under a constant velocity: d(v ln f)/dlnp = v x
    slope on the cells whose two faces are WENO3 faces."""
    E, M = 16, 3
    p = params(adiabaticChangeAlg=3)
    g = oracle.build_grids(p, E, M)
    lnp = g["dlnp"] * np.arange(E)
    lnf = np.stack([-3.0 - 2.5 * lnp] * M, axis=1)
    va = np.full((E + 1, M), v)
    f1 = oracle.change_operator(p, lnf, va, 7.0)
    np.testing.assert_allclose(f1[2:E - 2], v * -2.5, rtol=1e-12)
    # cells 1 and E-2: a WENO3 face (v ln f at the face, the alpha terms
    # cancelling) against the upwind face next to the boundary
    col = lnf[:, 0]
    up1 = col[0] if v > 0 else col[1]
    upE = col[E - 2] if v > 0 else col[E - 1]
    face2, faceE2 = 0.5 * (col[1] + col[2]), 0.5 * (col[E - 3] + col[E - 2])
    np.testing.assert_allclose(f1[1], v * (face2 - up1) / g["dlnp"], rtol=1e-12)
    np.testing.assert_allclose(f1[E - 2], v * (upE - faceE2) / g["dlnp"], rtol=1e-12)
    assert np.all(oracle.change_operator(p, lnf, np.zeros_like(va), 7.0) == 0.0)


def test_rk3_change_is_the_cubic_taylor_polynomial():
    """AdiabaticChange with alg 2 on a node whose change is affine in ln f
    (upwind operator + inflow boundary): each subcycle is the cubic Taylor
    polynomial of the affine map, checked through the whole step."""
    E, M = 7, 4
    p1, p2 = params(adiabaticChangeAlg=1), params(adiabaticChangeAlg=2)
    nodes = workloads.parker_nodes(1, 1, 10, seed=9, perturb=0.4, jitter=0.1)
    dims = workloads.dims_of(nodes, E, M)
    g = oracle.build_grids(p2, E, M)
    f0 = workloads.seeded_f(p2, g, nodes, noise=0.3, seed=10)
    rc, d2, f2, _, _, _ = oracle.update_ep(dims, p2, nodes, f0, op_flags=abi.EP_OP_CHANGE)
    assert rc == 0
    s, k = 1, 4
    dt = p2.tDel / p2.numEpSteps
    dt_full = p2.tDel
    dlnb, dlnn, dupar = nodes["dlnb"][s, k] / dt_full, nodes["dlnn"][s, k] / dt_full, nodes["dupar"][s, k] / dt_full
    damp = 1.0 - p2.dampCooling if dlnn < 0 else 1.0
    mu = g["mugrid"]
    vel = damp * ((-mu * dupar)[None, :] * g["vgrid_i"][:, None] +
                  (mu * mu * (dlnn - dlnb) + 0.5 * (1.0 - mu * mu) * dlnb)[None, :])
    va = np.zeros((E + 1, M))
    va[1:E] = 0.5 * (vel[1:] + vel[:-1])
    va[0], va[E] = va[1], va[E - 1]
    rmag = nodes["rmag"][s, k]
    c = oracle.change_operator(p2, np.zeros((E, M)), va, rmag).ravel()  # the inflow terms
    lin = probe_matrix(lambda x: oracle.change_operator(p2, x, va, rmag) - c.reshape(E, M), E, M)
    J = -lin  # d(ln f)/dt = -(lin ln f + c)
    n_sub = int(math.ceil(dt / (0.9 * g["dlnp"] / np.abs(vel).max())))
    h = dt / n_sub
    y = np.log(f0[s, k]).ravel()
    for _ in range(p2.numEpSteps):
        for _ in range(n_sub):
            r = J @ y - c
            y = y + h * r + h * h / 2 * (J @ r) + h ** 3 / 6 * (J @ (J @ r))
        y = np.log(np.maximum(np.exp(y), 2.2250738585072014e-308))
    np.testing.assert_allclose(f2[s, k].ravel(), np.exp(y), rtol=1e-11)


@pytest.mark.parametrize("v", [0.7, -0.45])
def test_rk3_upwind_focusing_is_total_variation_diminishing(v):
    """The Shu-Osher stages are convex combinations of forward-Euler upwind
    steps, so at a Courant number of 0.9 under a constant velocity the RK3
    upwind subcycle makes no new extrema and does not raise the total
    variation.  A compact profile away from the walls keeps the walls out
    of it.  A downwind operator breaks both."""
    M, dmu = 40, 2.0 / 40
    rng = np.random.default_rng(12)
    f = np.zeros((1, M))
    f[0, 14:26] = rng.uniform(1.0, 2.0, 12)
    va = np.full((1, M + 1), v)
    h = 0.9 * dmu / abs(v)

    def tv(x):
        return np.abs(np.diff(x[0])).sum()

    def subcycle(x, L):
        s1 = x - h * L(x)
        s1 = 0.75 * x + 0.25 * s1 - 0.25 * h * L(s1)
        return x / 3.0 + (2.0 / 3.0) * s1 - (2.0 / 3.0) * h * L(s1)

    up = lambda x: oracle.focus_operator(1, x, va, dmu)
    x = f.copy()
    for _ in range(6):
        y = subcycle(x, up)
        assert y.min() >= -1e-15 and y.max() <= x.max() + 1e-15
        assert tv(y) <= tv(x) * (1 + 1e-14)
        x = y
    assert np.sum(x) == pytest.approx(np.sum(f), rel=1e-13)
    # the plant: downwind differencing (the flux takes the downstream cell)
    down = lambda x: -oracle.focus_operator(1, x, -va, dmu)
    y = subcycle(f, down)
    assert y.min() < -1e-3 or tv(y) > tv(f) * 1.01

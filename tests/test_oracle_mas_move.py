"""Pins of the oracle's masMoveNodes (RK4 node motion through the MAS field,
src/masInterp.c:1279-1393) by closed forms:

* a uniform radial MAS wind moves every node straight out by
  V tDel / rScale: the four stage weights sum to one, the step is
  tDel / rScale and the velocity unit is MAS_V_CONVERT;
* an azimuthal wind V_phi = w r (solid rotation about z) in the equatorial
  plane makes the node ODE linear, dr/dt = a J r with J the quarter turn;
  the classical RK4 step of a linear ODE is the degree-4 Taylor polynomial
  of exp(a J), i.e. (1 - a^2/2 + a^4/24) I + (a - a^3/6) J, which an exact
  rotation, a midpoint or a Heun step all miss at the size of a used here;
* outside the MAS domain masWind carries the node's own previous state:
  its velocity scales by |r_old| / r of the node (so doubling r_old and
  halving V_phi moves the node the same way), and a radial previous wind
  moves it straight out;
* the node's position before the move becomes its old position, the inner
  node and the nodes past the active count do not move.
"""
import numpy as np
import pytest

import oracle
from eprem_amd import constants as K

RSCALE = 0.05


def level(vr=0.0, vp_per_rs=0.0, rmax=80.0):
    """One MAS level on exact FP32 meshes: r = 1..rmax Rs (integers),
    theta and phi covering the sphere; V_r = vr (MAS units) everywhere,
    V_phi = vp_per_rs * r; B and rho constant (unused by the motion)."""
    r = np.arange(1.0, rmax + 1.0, dtype=np.float32)
    t = np.linspace(0.05, 3.1, 12).astype(np.float32)
    p = np.linspace(0.0, 6.3, 16).astype(np.float32)
    P, T, R = np.meshgrid(p, t, r, indexing="ij")
    vals = {
        "bp": np.full(R.shape, 0.3), "bt": np.full(R.shape, -0.2), "br": np.full(R.shape, 1.5),
        "vp": vp_per_rs * R, "vt": np.zeros(R.shape), "vr": np.full(R.shape, vr), "rho": np.full(R.shape, 4.0),
    }
    fields = [vals[k].astype(np.float32).ravel() for k in ("bp", "bt", "br", "vp", "vt", "vr", "rho")]
    meshes = [m for _ in range(7) for m in (r, t, p)]
    return fields, meshes


def nodes_at(x, y, z):
    x, y, z = (np.asarray(a, np.float64).reshape(1, -1) for a in (x, y, z))
    return {"x": x.copy(), "y": y.copy(), "z": z.copy(), "rmag": np.sqrt(x * x + y * y + z * z)}


def state_for(nd, rold=None, **mhd):
    S, N = nd["x"].shape
    st = {k: np.zeros((S, N)) for k in oracle.MHD_NAMES}
    for k, v in mhd.items():
        st[k][:] = v
    src = nd if rold is None else rold
    st["rold_x"], st["rold_y"], st["rold_z"] = (np.array(src[k], np.float64) for k in ("x", "y", "z"))
    return st


def move(nd, st, tdel, lv, radial_max=60.0, vmin=0.0, n_active=None, first=1):
    return oracle.move_nodes_mas(nd, st, tdel, RSCALE, lv, lv, 0.0, 0.0, radial_max, vmin,
                                 n_active=n_active, first=first)


def test_uniform_radial_wind_moves_nodes_straight_out():
    rng = np.random.default_rng(1)
    n = 40
    r_rs = rng.uniform(5.0, 50.0, n)
    th, ph = rng.uniform(0.3, 2.8, n), rng.uniform(0.1, 6.2, n)
    ra = r_rs * K.RSAU / RSCALE
    x, y, z = ra * np.sin(th) * np.cos(ph), ra * np.sin(th) * np.sin(ph), ra * np.cos(th)
    nd = nodes_at(np.r_[0.0 + x[0], x], np.r_[y[0], y], np.r_[z[0], z])
    r0 = nd["rmag"].copy()
    dirs = np.stack([nd[k] / nd["rmag"] for k in ("x", "y", "z")])
    vr_mas, tdel = 0.75, 3.0
    move(nd, state_for(nd), tdel, level(vr=vr_mas))
    step = vr_mas * K.MAS_V_CONVERT * tdel / RSCALE
    np.testing.assert_allclose(nd["rmag"][0, 1:], r0[0, 1:] + step, rtol=1e-13)
    assert nd["rmag"][0, 0] == r0[0, 0]  # the inner node stays
    new_dirs = np.stack([nd[k] / nd["rmag"] for k in ("x", "y", "z")])
    np.testing.assert_allclose(new_dirs, dirs, atol=1e-14)


def test_solid_rotation_step_is_the_quartic_taylor_polynomial():
    rng = np.random.default_rng(2)
    n = 30
    r_rs = rng.uniform(5.0, 50.0, n)
    ph = rng.uniform(0.0, 2 * np.pi, n)
    ra = r_rs * K.RSAU / RSCALE
    x0, y0 = ra * np.cos(ph), ra * np.sin(ph)
    nd = nodes_at(np.r_[1.0, x0], np.r_[0.5, y0], np.zeros(n + 1))
    w, tdel = 2.0 ** -6, 50.0
    a = tdel * w * K.MAS_V_CONVERT / K.RSAU
    assert 0.2 < a < 0.4
    old = move(nd, state_for(nd), tdel, level(vp_per_rs=w))
    c, s = 1 - a * a / 2 + a ** 4 / 24, a - a ** 3 / 6
    np.testing.assert_allclose(nd["x"][0, 1:], c * x0 - s * y0, rtol=1e-12, atol=1e-12 * ra.max())
    np.testing.assert_allclose(nd["y"][0, 1:], s * x0 + c * y0, rtol=1e-12, atol=1e-12 * ra.max())
    assert np.all(nd["z"] == 0.0)
    # RK4 is not the exact rotation, nor a second-order step
    rot = np.cos(a) * x0 - np.sin(a) * y0
    mid = (1 - a * a / 2) * x0 - a * y0
    for other in (rot, mid):
        assert np.max(np.abs(nd["x"][0, 1:] - other)) > 1e-7 * ra.max()
    np.testing.assert_array_equal(old["rold_x"][0, 1:], x0)
    np.testing.assert_array_equal(old["rold_y"][0, 1:], y0)


def test_outside_the_domain_the_node_carries_its_own_wind():
    rng = np.random.default_rng(3)
    n = 20
    r_rs = rng.uniform(70.0, 200.0, n)  # beyond mas_radial_max = 60 Rs
    ph = rng.uniform(0.0, 2 * np.pi, n)
    ra = r_rs * K.RSAU / RSCALE
    x0, y0 = ra * np.cos(ph), ra * np.sin(ph)
    lv = level(vr=5.0, vp_per_rs=1.0)  # inside values, never used out here
    tdel = 20.0
    # radial previous wind: straight out by vr tdel / rScale
    nd = nodes_at(np.r_[1.0, x0], np.r_[0.0, y0], np.zeros(n + 1))
    vr = 1.3e-3
    move(nd, state_for(nd, vr=vr, vtheta=0.0, vphi=0.0), tdel, lv)
    np.testing.assert_allclose(nd["rmag"][0, 1:], ra + vr * tdel / RSCALE, rtol=1e-13)
    # azimuthal previous wind scaled by |r_old| / r of the node
    runs = []
    for fac, vp in ((1.0, 4.0e-4), (2.0, 2.0e-4), (0.5, 8.0e-4)):
        nd = nodes_at(np.r_[1.0, x0], np.r_[0.0, y0], np.zeros(n + 1))
        rold = {k: nd[k] * fac for k in ("x", "y", "z")}
        move(nd, state_for(nd, rold=rold, vr=2.0e-4, vphi=vp), tdel, lv)
        runs.append(np.stack([nd["x"], nd["y"]]))
    np.testing.assert_allclose(runs[1], runs[0], rtol=1e-13)
    np.testing.assert_allclose(runs[2], runs[0], rtol=1e-13)
    # and the move is not radial: the azimuthal part acted
    dphi = np.angle(np.exp(1j * (np.arctan2(runs[0][1, 0, 1:], runs[0][0, 0, 1:]) - ph)))
    assert np.all(dphi > 0)


def test_inactive_and_inner_nodes_stay():
    lv = level(vr=1.0)
    ra = np.array([10.0, 20.0, 30.0, 40.0, 50.0]) * K.RSAU / RSCALE
    x = np.stack([ra, ra])
    nd = {"x": x.copy(), "y": np.zeros_like(x), "z": np.zeros_like(x), "rmag": x.copy()}
    before = {k: v.copy() for k, v in nd.items()}
    st = state_for(nd)
    old = oracle.move_nodes_mas(nd, st, 2.0, RSCALE, lv, lv, 0.0, 0.0, 60.0, 0.0, n_active=[5, 3])
    assert np.all(nd["x"][0, 1:] > before["x"][0, 1:]) and np.all(nd["x"][1, 1:3] > before["x"][1, 1:3])
    np.testing.assert_array_equal(nd["x"][:, 0], before["x"][:, 0])
    np.testing.assert_array_equal(nd["x"][1, 3:], before["x"][1, 3:])
    np.testing.assert_array_equal(old["rold_x"][:, 0], before["x"][:, 0])
    np.testing.assert_array_equal(old["rold_x"][0, 1:], before["x"][0, 1:])

"""Pins of the oracle's MAS path (masTriLinear, masWind; src/masInterp.c:
30-68, :117-200, :539-741, :1007-1065) by what the physics and the
mathematics fix, independently of the oracle's own formulas:

* every one of the seven field slots lands in its own output component,
  with the unit conversion of its kind stated from the MAS normalisation
  (velocity unit R_sun / 1445.87 s = 481.37 km/s, density unit
  1e8 protons / cm^3) and, for B, from the Alfven speed, which must come
  out the same in MAS units and in code units;
* trilinear interpolation is exact for multilinear fields, also on the
  reference's own staggered MAS meshes (tests/golden/mas_grids, read from
  testsuite/mas_*/input/data/*.h5) and returns the nodal values at mesh
  nodes of the real mas_wind data;
* outside masRadialMax a radially moving node carries a steady radial wind
  exactly: B_r and density fall as r^-2, the transverse B and V components
  as r^-1, V_r is constant;
* V_r is clamped at masVmin / c;
* phi is shifted by -phiOffset and wrapped into [0, 2 pi].
Each check fails under a swapped slot, a wrong exponent or a wrong sign.
"""
import math
import os

import numpy as np
import pytest

import oracle
from eprem_amd import constants as K

GRIDS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mas_grids")
FIELDS = ("bp", "bt", "br", "vp", "vt", "vr", "rho")
OUT = ("bphi", "btheta", "br", "vphi", "vtheta", "vr", "density")  # output of slot k

# MAS units stated from their definition, not from the oracle's macros.
V_UNIT_KMS = 481.3711  # R_sun / t_MAS = 6.96e5 km / 1445.87 s
V_UNIT = V_UNIT_KMS * 1e5 / 2.99792458e10  # in c
N_UNIT = 1.0e8  # rho_MAS = 1.6726e-16 g cm^-3 = 1e8 proton masses per cm^3
# Alfven speed v_A = B_MAS / sqrt(rho_MAS) in MAS velocity units; in code
# units (v_A / c)^2 = B^2 / (4 pi n) (src/baseTypes.h:89-95).  Equating both
# for any cell gives the B conversion: B_code = B_MAS * V_UNIT * sqrt(4 pi N_UNIT).
B_UNIT = V_UNIT * math.sqrt(4.0 * math.pi * N_UNIT)
# The checks below use the reference's constants (src/baseTypes.h:108-113,
# restated in eprem_amd.constants), which test_units_match_the_mas_
# normalisation holds to the definitions above at their 7 stated digits.
UNIT = (K.MAS_B_CONVERT,) * 3 + (K.MAS_V_CONVERT,) * 3 + (K.MAS_RHO_CONVERT,)

RSCALE = 0.01


def cart(r_rs, th, ph, rscale=RSCALE):
    """Node position (rScale units) of a point at r_rs solar radii."""
    rau = np.asarray(r_rs) * K.RSAU
    x = rau * np.sin(th) * np.cos(ph) / rscale
    y = rau * np.sin(th) * np.sin(ph) / rscale
    z = rau * np.cos(th) / rscale
    return x, y, z, np.sqrt(x * x + y * y + z * z)


def load(name):
    z = np.load(os.path.join(GRIDS, name))
    return {k: z[k] for k in z.files}


def real_meshes(z, level=1):
    return [z["%d_%s_%s" % (level, f, a)] for f in FIELDS for a in "rtp"]


# seven different multilinear fields (coefficients per slot)
COEF = [(2.0, 0.3, -0.5, 0.1, 0.02, -0.01, 0.04, 0.003),
        (-1.5, 0.2, 0.7, -0.2, 0.01, 0.02, -0.03, 0.001),
        (3.0, -0.1, 0.4, 0.3, -0.02, 0.01, 0.05, -0.002),
        (0.5, 0.15, -0.3, -0.1, 0.03, -0.02, 0.02, 0.004),
        (-0.8, -0.05, 0.6, 0.25, 0.01, 0.03, -0.04, -0.001),
        (4.0, 0.25, 0.2, -0.15, -0.01, 0.015, 0.03, 0.002),
        (1.0, 0.4, -0.6, 0.2, 0.025, -0.015, -0.02, 0.0025)]


def ml(k, r, t, p):
    a = COEF[k]
    return a[0] + a[1] * r + a[2] * t + a[3] * p + a[4] * r * t + a[5] * r * p + a[6] * t * p + a[7] * r * t * p


def ml_scale(k, r, t, p):
    """Sum of the terms' magnitudes: the scale of the FP32 rounding error."""
    a = COEF[k]
    return sum(abs(c) * abs(v) for c, v in zip(a, (1, r, t, p, r * t, r * p, t * p, r * t * p)))


def sample(k, meshes):
    r, t, p = (m.astype(np.float64) for m in meshes[3 * k:3 * k + 3])
    P, T, R = np.meshgrid(p, t, r, indexing="ij")  # [phi][theta][r], r fastest
    return ml(k, R, T, P).astype(np.float32).ravel()


def run(meshes, fields, r_rs, th, ph, s=0.0, phi_offset=0.0, radial_max=1.0e4, vmin=0.0, fields1=None,
        prev=None, rscale=RSCALE):
    x, y, z, rmag = cart(r_rs, th, ph, rscale)
    lv0 = (fields, meshes)
    lv1 = (fields1 if fields1 is not None else fields, meshes)
    return oracle.mas_interp(rscale, x, y, z, rmag, lv0, lv1, s, phi_offset, radial_max, vmin, prev)


def query(n, seed, rlo=2.0, rhi=200.0, tlo=0.2, thi=2.9, plo=0.2, phi=6.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(rlo, rhi, n), rng.uniform(tlo, thi, n), rng.uniform(plo, phi, n)


@pytest.mark.parametrize("grid", ["mas_wind_levels.npz", "mas_cme_meshes.npz", "synthetic"])
def test_seven_slots_units_and_staggered_meshes(grid):
    """Distinct multilinear fields in the seven slots, each on its own
    staggered mesh: every output is its own field times its unit."""
    if grid == "synthetic":
        rng = np.random.default_rng(3)
        meshes = []
        for _ in range(7):
            meshes += [np.sort(np.r_[1.0, rng.uniform(1.0, 220.0, 15), 220.0]).astype(np.float32),
                       np.sort(np.r_[-0.1, rng.uniform(-0.1, 3.3, 11), 3.3]).astype(np.float32),
                       np.sort(np.r_[-0.1, rng.uniform(-0.1, 6.4, 9), 6.4]).astype(np.float32)]
        rhi = 200.0
    else:
        meshes = real_meshes(load(grid))
        rhi = 17.0 if "cme" in grid else 200.0
    fields = [sample(k, meshes) for k in range(7)]
    r, t, p = query(40, 5, rhi=rhi)
    out = run(meshes, fields, r, t, p)
    for k in range(7):
        exact = ml(k, r, t, p) * UNIT[k]
        tol = 3e-7 * ml_scale(k, r, t, p) * UNIT[k]
        err = np.abs(out[OUT[k]] - exact)
        assert np.all(err <= tol), (FIELDS[k], float(np.max(err / tol)))


def test_units_match_the_mas_normalisation():
    """The conversions equal the MAS units stated from their definition,
    and differ by orders of magnitude, so the slot checks catch a field
    given the wrong kind of conversion."""
    assert B_UNIT / V_UNIT > 1e4 and N_UNIT / B_UNIT > 1e5
    assert K.MAS_V_CONVERT == pytest.approx(V_UNIT, rel=1e-6)
    assert K.MAS_B_CONVERT == pytest.approx(B_UNIT, rel=1e-6)
    assert K.MAS_RHO_CONVERT == pytest.approx(N_UNIT, rel=1e-12)


def test_time_blend_of_two_levels():
    """(1 - s) level0 + s level1 for every slot."""
    z = load("mas_wind_levels.npz")
    meshes = real_meshes(z)
    f0 = [sample(k, meshes) for k in range(7)]
    f1 = [(2.0 * a - 0.5).astype(np.float32) for a in f0]
    r, t, p = query(20, 8)
    s = 0.3
    out = run(meshes, f0, r, t, p, s=s, fields1=f1)
    for k in range(7):
        v0 = ml(k, r, t, p)
        exact = ((1 - s) * v0 + s * (2.0 * v0 - 0.5)) * UNIT[k]
        tol = 1e-6 * (ml_scale(k, r, t, p) + 1.0) * UNIT[k]
        assert np.all(np.abs(out[OUT[k]] - exact) <= tol), FIELDS[k]


def test_nodal_values_of_the_real_mas_wind_data():
    """At a mesh node of a field, the interpolant is that node's value
    (converted); the query is built from the node's (r, theta, phi) via
    Cartesian coordinates, as a real node position is."""
    z = load("mas_wind_levels.npz")
    meshes = real_meshes(z)
    fields = [z["1_%s_data" % f] for f in FIELDS]
    rng = np.random.default_rng(4)
    for k in range(7):
        rm, tm, pm = (m.astype(np.float64) for m in meshes[3 * k:3 * k + 3])
        nr, nt = len(rm), len(tm)
        # interior nodes with theta inside (0, pi) and phi inside (0, 2 pi)
        it = [j for j in range(nt) if 0.05 < tm[j] < math.pi - 0.05]
        ip = [j for j in range(len(pm)) if 0.05 < pm[j] < 2 * math.pi - 0.05]
        ir = rng.integers(5, nr - 5, 6)
        for i in ir:
            for j in it:
                for l in ip:
                    out = run(meshes, fields, np.array([rm[i]]), np.array([tm[j]]), np.array([pm[l]]))
                    nodal = float(fields[k][i + nr * j + nr * nt * l]) * UNIT[k]
                    if k == 5:
                        nodal = max(nodal, 0.0)  # masVmin = 0 clamps a negative V_r
                    # neighbouring nodes differ by O(1): the position's
                    # rounding moves the weight by ~1e-13
                    assert out[OUT[k]][0] == pytest.approx(nodal, rel=1e-9, abs=1e-12 * UNIT[k]), \
                        (FIELDS[k], i, j, l)


def test_vmin_clamp():
    """V_r below masVmin / c is raised to it; other slots are untouched."""
    z = load("mas_wind_levels.npz")
    meshes = real_meshes(z)
    fields = [sample(k, meshes) for k in range(7)]
    rr, tt, pp = (m.astype(np.float64) for m in meshes[15:18])
    P, T, R = np.meshgrid(pp, tt, rr, indexing="ij")
    fields[5] = (0.05 + 0.01 * R).astype(np.float32).ravel()  # V_r in MAS units
    vmin_kms = 50.0  # 0.05 + 0.01 r crosses 50 km/s at r = 5.39 Rs
    r = np.array([2.0, 4.0, 10.0, 20.0, 100.0])
    t, p = np.full(5, 1.3), np.full(5, 2.2)
    out = run(meshes, fields, r, t, p, vmin=vmin_kms * 1e5)
    floor = vmin_kms * 1e5 / 2.99792458e10
    raw = (0.05 + 0.01 * r) * V_UNIT
    assert raw[0] < floor and raw[1] < floor and np.all(raw[2:] > floor)
    np.testing.assert_array_equal(out["vr"][:2], [floor, floor])
    np.testing.assert_allclose(out["vr"][2:], raw[2:], rtol=1e-6)
    np.testing.assert_allclose(out["density"], ml(6, r, t, p) * N_UNIT, rtol=1e-6)


@pytest.mark.parametrize("phi_node,offset", [(0.3, 1.0), (-2.0, -2.5), (2.0, 0.5), (5.9, -0.2),
                                             (1.0, 2 * math.pi + 0.4), (4.0, -2 * math.pi - 0.7)])
def test_phi_offset_shift_and_wrap(phi_node, offset):
    """The field is sampled at phi_node - phiOffset taken into [0, 2 pi]."""
    rm = np.array([1.0, 5.0, 50.0], np.float32)
    tm = np.array([0.0, 1.0, 2.0, 3.2], np.float32)
    pm = np.linspace(-0.5, 2 * math.pi + 0.5, 40).astype(np.float32)
    P, T, R = np.meshgrid(pm.astype(float), tm.astype(float), rm.astype(float), indexing="ij")
    fld = (1.0 + 0.1 * P).astype(np.float32).ravel()
    meshes = [rm, tm, pm] * 7
    out = run(meshes, [fld] * 7, np.array([10.0]), np.array([1.4]), np.array([phi_node]),
              phi_offset=offset)
    eff = (phi_node - offset) % (2 * math.pi)
    expect = (1.0 + 0.1 * eff) * N_UNIT
    assert out["density"][0] == pytest.approx(expect, rel=2e-6)
    # the opposite sign of the offset lands elsewhere
    wrong = (1.0 + 0.1 * ((phi_node + offset) % (2 * math.pi))) * N_UNIT
    assert abs(wrong - expect) > 1e-3 * expect


def parker(r_au, th, b0=3.0, v=1.3e-3, om=2.9e-3, d0=7.0, vt0=2e-5, vp0=4e-5, bt0=0.4):
    """Steady radial wind at constant speed v from a reference sphere at
    1 AU: flux conservation gives B_r ~ r^-2, frozen-in transverse field
    B_phi = -B_r om r sin(theta) / v ~ r^-1 (Parker) and B_theta ~ r^-1,
    mass conservation n ~ r^-2, angular momentum v_phi, v_theta ~ r^-1."""
    br = b0 / r_au ** 2
    return dict(br=br, btheta=bt0 / r_au, bphi=-br * om * r_au * np.sin(th) / v, vr=np.full_like(r_au, v),
                vtheta=vt0 / r_au, vphi=vp0 / r_au, density=d0 / r_au ** 2)


def test_maswind_carries_a_steady_radial_wind():
    """Outside masRadialMax a node that moved radially from r_old to r
    gets the wind's exact values at r from its values at r_old."""
    rng = np.random.default_rng(9)
    n = 30
    th, ph = rng.uniform(0.2, 2.9, n), rng.uniform(0.0, 6.2, n)
    r_old_au = rng.uniform(0.2, 3.0, n)
    r_au = r_old_au * rng.uniform(1.0001, 1.3, n)
    rscale = 0.00465047
    x, y, z, rmag = cart(r_au / K.RSAU, th, ph, rscale)
    ox, oy, oz, _ = cart(r_old_au / K.RSAU, th, ph, rscale)
    old, new = parker(r_old_au, th), parker(r_au, th)
    prev = dict(rold_x=ox, rold_y=oy, rold_z=oz, **old)
    meshes = [np.array([1.0, 2.0], np.float32), np.array([0.0, 3.2], np.float32),
              np.array([0.0, 6.3], np.float32)] * 7
    fld = np.ones(8, np.float32)
    out = oracle.mas_interp(rscale, x, y, z, rmag, ([fld] * 7, meshes), ([fld] * 7, meshes), 0.5, 0.0,
                            30.0, 0.0, prev)
    assert np.all(r_au > 30.0 * K.RSAU)
    for k in OUT:
        np.testing.assert_allclose(out[k], new[k], rtol=1e-13, atol=0, err_msg=k)

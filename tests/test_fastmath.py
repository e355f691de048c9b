"""The table-driven exp and log of the fused precision modes against
extended-precision references (numpy longdouble: x87 80-bit), on the host
build of the same code the device runs; a GPU test checks that the device
gives the same bits.

Bars: exp within 1 ulp (relative 2^-52) wherever the result is a normal
double; log within 2^-52 * max(1, |ln x|) absolute -- the adiabatic change
advects ln f, so an absolute error e of ln f is a relative error e of f."""
import numpy as np
import pytest

from eprem_amd import runtime

RNG = np.random.default_rng(20261203)


def _log_inputs():
    return np.concatenate([np.exp(RNG.uniform(-708.0, 709.0, 100000)), 1.0 + RNG.uniform(-1e-3, 1e-3, 20000),
                           RNG.uniform(0.5, 2.0, 20000), 1.0 + np.arange(-64, 65) / 128.0,
                           np.array([1.0, 2.0, 0.5, 2.2250738585072014e-308, 1.7976931348623157e308,
                                     1.1754943508222875e-38])])


def _exp_inputs():
    return np.concatenate([RNG.uniform(-708.0, 709.0, 100000), RNG.uniform(-1.0, 1.0, 20000),
                           np.arange(-64, 65) * (np.log(2.0) / 64.0), np.array([0.0, -1e-300, 1e-300, 709.7])])


def test_log_accuracy():
    x = _log_inputs()
    lg, _ = runtime.fastmath(x)
    ref = np.log(x.astype(np.longdouble))
    err = np.abs(lg.astype(np.longdouble) - ref)
    bound = np.maximum(1.0, np.abs(ref)) * np.longdouble(2.0 ** -52)
    assert np.all(err <= bound), float(np.max(err / bound))


def test_exp_accuracy():
    y = _exp_inputs()
    _, ex = runtime.fastmath(y)
    ref = np.exp(y.astype(np.longdouble))
    normal = ref >= np.longdouble(2.2250738585072014e-308)
    rel = np.abs(ex[normal].astype(np.longdouble) - ref[normal]) / ref[normal]
    assert np.all(rel <= np.longdouble(2.0 ** -52)), float(np.max(rel))


def test_special_values():
    lg, ex = runtime.fastmath(np.array([np.inf, np.nan, 710.0, -746.0, -np.inf]))
    assert lg[0] == np.inf and np.isnan(lg[1])
    assert ex[0] == np.inf and np.isnan(ex[1]) and ex[2] == np.inf and ex[3] == 0.0 and ex[4] == 0.0


def test_exp_log_round_trip():
    """exp(log(f)) returns f over the 35 decades the distribution spans
    (the change operator's first and last step with no flux) to the
    accuracy a double ln f allows: 2^-52 (|ln f| + 2)."""
    f = 10.0 ** RNG.uniform(-30.0, 5.0, 50000)
    lg, _ = runtime.fastmath(f)
    _, back = runtime.fastmath(lg)
    bound = 2.0 ** -52 * (np.maximum(1.0, np.abs(np.log(f))) + 2.0)
    assert np.all(np.abs(back - f) / f <= bound)


@pytest.mark.gpu
def test_device_bits_equal_host():
    from eprem_amd import config, workloads
    x = np.concatenate([_log_inputs()[:40000], np.abs(_exp_inputs()[:40000]) + 1e-300])
    y = _exp_inputs()[:80000]
    nodes = workloads.parker_nodes(1, 1, 8, seed=1)
    sol = runtime.EpSolver(workloads.dims_of(nodes, 4, 4), config.default_params())
    try:
        dl, _ = runtime.fastmath(x, device=True)
        _, de = runtime.fastmath(y, device=True)
    finally:
        sol.close()
    hl, _ = runtime.fastmath(x)
    _, he = runtime.fastmath(y)
    np.testing.assert_array_equal(dl, hl)
    np.testing.assert_array_equal(de, he)

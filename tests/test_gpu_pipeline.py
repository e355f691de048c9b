"""Whole analytic-wind runs of the reference's own test cases (seed, wind)
with node spawn, motion, updateMhd and the EP update on the device: the
observer files the run writes must digest to the reference's committed
statistics, and the distribution must match the oracle's run."""
import os

import numpy as np
import pytest

import oracle
from eprem_amd import config, digest, driver

pytestmark = pytest.mark.gpu


def _case(case, golden_dir, tmp_path):
    cfg = config.read_config(open(os.path.join(golden_dir, case, "eprem_input_file")).read())
    run = driver.ParkerRun(cfg)
    try:
        res = run.run(outdir=str(tmp_path))
    finally:
        run.close()
    ref = digest.parse_digest(open(os.path.join(golden_dir, case, "eprem_digest.txt")).read())
    mine = digest.parse_digest(digest.digest_dir(str(tmp_path)))
    assert sorted(mine) == sorted(ref)
    stats_bad = [(fn, v) for fn in ref for v in ref[fn]
                 if mine[fn][v]["shape"] != ref[fn][v]["shape"] or mine[fn][v]["stats"] != ref[fn][v]["stats"]]
    ores = oracle.run_parker(cfg, nthreads=8, raw=True)
    assert ores["rc"] == 0
    return cfg, res, ores, stats_bad


def test_seed_run_on_device(golden_dir, tmp_path):
    cfg, res, ores, stats_bad = _case("seed", golden_dir, tmp_path)
    assert stats_bad == []
    np.testing.assert_allclose(res["dist"], ores["dist"], rtol=1e-12, atol=0)
    np.testing.assert_array_equal(res["time"], ores["time"])


def test_wind_run_on_device(golden_dir, tmp_path):
    cfg, res, ores, stats_bad = _case("wind", golden_dir, tmp_path)
    assert stats_bad == []
    assert res["dist"].shape == ores["dist"].shape
    rel = np.abs(res["dist"] - ores["dist"]) / np.abs(ores["dist"])
    assert float(rel.max()) <= 1e-10
    # per-step subcycle maxima equal the oracle's (and so the reference log's)
    od = ores["diag"]
    assert len(res["diag"]) == len(od)
    for a, b in zip(res["diag"], od):
        assert a["max_subcycles_focusing"] == b["max_subcycles_focusing"]
        assert a["max_subcycles_change"] == b["max_subcycles_change"]

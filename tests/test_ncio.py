"""The observer-file writer against the reference's own digests: the
oracle's seed and wind runs are written in the reference's per-stream file
layout, read back with an independent netCDF reader (scipy) and digested;
the digests must equal the ones committed with the reference's tests."""
import os

import numpy as np
import pytest

import oracle
from eprem_amd import config, digest, ncio
from scipy.io import netcdf_file


def _write_run(cfg, res, outdir):
    S = config.stream_count(cfg)
    for s in range(S):
        raw = {k: res[k][:, s] for k in ("R", "T", "P", "Br", "Bt", "Bp", "Vr", "Vt", "Vp", "Rho")}
        raw["Dist"] = res["dist"][:, s][:, :, None, :, :]
        raw.update(time=res["time"], mu=res["mu"], egrid=res["egrid"], vgrid=res["vgrid"],
                   mass=np.asarray(cfg["mass"], dtype=np.float64), charge=np.asarray(cfg["charge"], dtype=np.float64))
        ncio.write_observer_file(os.path.join(outdir, ncio.stream_file_name(s)), raw, cfg["rScale"])


def _check(case, golden_dir, tmp_path, max_sha_mismatch):
    cfg = config.read_config(open(os.path.join(golden_dir, case, "eprem_input_file")).read())
    res = oracle.run_parker(cfg, nthreads=8, raw=True)
    assert res["rc"] == 0
    _write_run(cfg, res, str(tmp_path))
    mine = digest.parse_digest(digest.digest_dir(str(tmp_path)))
    ref = digest.parse_digest(open(os.path.join(golden_dir, case, "eprem_digest.txt")).read())
    assert sorted(mine) == sorted(ref)
    sha_bad = []
    for fn in ref:
        assert sorted(mine[fn]) == sorted(ref[fn]), fn
        for v in ref[fn]:
            assert mine[fn][v]["shape"] == ref[fn][v]["shape"], (fn, v)
            assert mine[fn][v]["stats"] == ref[fn][v]["stats"], (fn, v)
            if mine[fn][v]["sha256"] != ref[fn][v]["sha256"]:
                sha_bad.append((fn, v))
    assert len(sha_bad) <= max_sha_mismatch, sha_bad
    return tmp_path


def test_seed_files_digest_bit_exact(golden_dir, tmp_path):
    _check("seed", golden_dir, tmp_path, 0)


def test_wind_files_digest(golden_dir, tmp_path):
    # same 18 position lines as the in-memory comparison (azimuth libm
    # rounding on three streams); everything else bit for bit
    _check("wind", golden_dir, tmp_path, 18)


def test_file_layout(golden_dir, tmp_path):
    cfg = config.read_config(open(os.path.join(golden_dir, "seed", "eprem_input_file")).read())
    res = oracle.run_parker(cfg, nthreads=8, raw=True)
    _write_run(cfg, res, str(tmp_path))
    with netcdf_file(str(tmp_path / "obs000000.nc"), "r", mmap=False) as ds:
        assert ds.dimensions["time"] is None  # record dimension
        assert ds.variables["Dist"].dimensions == ("time", "shell", "species", "energy", "mu")
        assert ds.variables["R"].units == b"au"
        assert ds.variables["Dist"].scale_factor == pytest.approx(1.0 / 27.0, rel=0, abs=0)
        assert ds.variables["shell"].data.dtype.kind == "i"

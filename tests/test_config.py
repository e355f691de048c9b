"""Input-file reader (the libconfig subset EPREM uses) and the code-unit
conversion of setRuntimeConstants (src/configuration.c:324-348)."""
import os

import pytest

from eprem_amd import config, constants as K

CASES = ("seed", "wind", "cone", "mas_wind", "mas_cme")


@pytest.mark.parametrize("case", CASES)
def test_reads_reference_inputs(case, golden_dir):
    cfg = config.read_config(open(os.path.join(golden_dir, case, "eprem_input_file")).read())
    assert cfg["_type_mismatch"] == []
    assert cfg["numSpecies"] == 1


def test_wind_values(golden_dir):
    cfg = config.read_config(open(os.path.join(golden_dir, "wind", "eprem_input_file")).read())
    assert cfg["numNodesPerStream"] == 60 and cfg["numEnergySteps"] == 12 and cfg["numMuSteps"] == 6
    assert cfg["tDel"] == 0.041666666666667 and cfg["numEpSteps"] == 5
    assert cfg["mfpRadialPower"] == 2.0  # default, as echoed in the reference run params
    p = config.ep_params(cfg)
    assert p.tDel == 0.041666666666667 / K.DAY
    assert p.mhdBsAu == 1.60e-5 / K.MHD_B_NORM


def test_type_mismatch_falls_back_to_default():
    """libconfig's lookup_float fails on an integer literal, so the
    reference silently keeps the default value."""
    cfg = config.read_config("tDel=1\nnumEpSteps=7;\n")
    assert cfg["tDel"] == 0.01041666666667
    assert cfg["numEpSteps"] == 7
    assert "tDel" in cfg["_type_mismatch"]


def test_range_check_panics():
    with pytest.raises(config.ConfigError):
        config.read_config("numMuSteps=1\n")
    with pytest.raises(config.ConfigError):
        config.read_config("dampCooling=1.5\n")


def test_comments_arrays_strings():
    raw = config.parse_libconfig('# c\nmass=[1.0, 4.0];  // x\n/* block\n */ masDirectory="a b"\nx=1e-3\n')
    assert raw == {"mass": [1.0, 4.0], "masDirectory": "a b", "x": 1e-3}


@pytest.mark.parametrize("case,bad", [("cone", ["idealShock"]), ("mas_wind", ["masCouple"]),
                                      ("mas_cme", ["masCouple"])])
def test_host_loop_rejects_what_it_does_not_run(case, bad, golden_dir):
    """The analytic-wind host loop refuses inputs it would run as a plain
    Parker wind: the cone test's ideal shock and the MAS-coupled cases
    (the check runs before any device call, so this needs no GPU)."""
    from eprem_amd import driver
    cfg = config.read_config(open(os.path.join(golden_dir, case, "eprem_input_file")).read())
    with pytest.raises(config.ConfigError) as ei:
        driver.ParkerRun(cfg)
    for k in bad:
        assert k in str(ei.value)


@pytest.mark.parametrize("opt", config.UNSUPPORTED_BY_HOST_LOOP)
def test_host_loop_rejects_each_option(opt, golden_dir):
    cfg = config.read_config(open(os.path.join(golden_dir, "wind", "eprem_input_file")).read())
    config.check_host_loop_options(cfg)  # the wind case itself is supported
    cfg[opt] = 1
    with pytest.raises(config.ConfigError, match=opt):
        config.check_host_loop_options(cfg)

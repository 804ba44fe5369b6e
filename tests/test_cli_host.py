"""Command-line driver: flags, defaults and derived parameters (host only)."""
import numpy as np

import oracle as O
from pyminiweather_amd.__main__ import derived_params, get_params_from_args, get_parser
from pyminiweather_amd.output import svars_filename

# pyminiweather/__main__.py:28-156: dest -> default
REFERENCE_DEFAULTS = {"nx": 200, "nz": 100, "xlen": 2e4, "zlen": 1e4, "nsteps": 10, "dt": None,
                      "nwarmups": 0, "ic_type": "thermal", "hs": 2, "s": 4,
                      "max_speed": 500.0, "cfl": 1.0, "verbose": False, "output_freq": -1,
                      "app_filename": "PyMiniWeatherData.txt", "app_log_file": None}


def test_defaults_and_params_dict_match_reference():
    args, _ = get_parser().parse_known_args([])
    params = get_params_from_args(args)
    assert params == REFERENCE_DEFAULTS


def test_ic_choices_and_flag_spellings():
    parser = get_parser()
    for ic in ("thermal", "collision", "gravity", "density-current", "injection"):
        assert parser.parse_args(["--ic-type", ic]).ic_type == ic
    a = parser.parse_args(["--max-speed", "450", "--output-freq", "3", "--app-filename",
                           "a.txt", "--nwarmups", "2", "--cfl", "1.5", "--verbose"])
    assert (a.max_speed, a.output_freq, a.app_filename, a.nwarmups, a.cfl, a.verbose) == \
        (450.0, 3, "a.txt", 2, 1.5, True)
    # unknown flags are ignored, as with parse_known_args in the reference
    a, rest = parser.parse_known_args(["--nx", "8", "--bogus", "1"])
    assert a.nx == 8 and rest == ["--bogus", "1"]


def test_derived_dt_equals_reference_formula():
    for nx, nz, cfl, ms in ((100, 50, 1.5, 450.0), (2000, 1000, 1.0, 500.0), (37, 11, 0.8, 300.0)):
        args, _ = get_parser().parse_known_args(["--nx", str(nx), "--nz", str(nz), "--cfl",
                                                 str(cfl), "--max-speed", str(ms)])
        p = derived_params(get_params_from_args(args))
        ref = O.make_params(nx, nz, cfl=cfl, max_speed=ms)
        assert p["dx"] == ref["dx"] and p["dz"] == ref["dz"] and p["dt"] == ref["dt"]
        assert isinstance(p["dt"], np.floating)  # np.minimum result, as in the reference
    args, _ = get_parser().parse_known_args(["--dt", "0.25"])
    assert derived_params(get_params_from_args(args))["dt"] == 0.25


def test_svars_filename():
    assert svars_filename("PyMiniWeatherData.txt") == "PyMiniWeatherData_svars.txt"
    assert svars_filename("out.csv") == "out_svars.csv"

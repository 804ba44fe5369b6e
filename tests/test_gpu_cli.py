"""The command-line driver end to end on the GPU, against the oracle."""
import re

import numpy as np
import pytest

import oracle as O
from conftest import load_golden, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ic,warmups", [("collision", 0), ("injection", 0),
                                        ("density-current", 2)])
def test_cli_outputs_and_log_match_reference(tmp_path, ic, warmups):
    """40x20, 5 steps (after `warmups` warm-up steps), output every 2 steps:
    the state and solution-variable files hold the snapshots taken before
    steps 1 and 3 in the reference's savetxt layout, and the log reports the
    start (before the warm-ups) and end mass and energy and max(rho*theta).
    Compared with the reference's own driver loop and
    compute_solution_variables (golden ``cli_*``, pyminiweather/__main__.py:
    205-247, post/stats.py:38-69). The injection IC runs the per-stage
    kernels."""
    import pyminiweather_amd.solve as S
    from pyminiweather_amd.__main__ import main
    g = load_golden(f"cli_{ic}_40x20")
    assert int(g["nwarmups"]) == warmups and int(g["output_freq"]) == 2
    out = tmp_path / "run.txt"
    log = tmp_path / "log.txt"
    S.reverse_direction = False
    rc = main(["--nx", "40", "--nz", "20", "--nsteps", "5", "--output-freq", "2",
               "--nwarmups", str(warmups), "--ic-type", ic, "--app-filename", str(out),
               "--app-log-file", str(log)])
    S.reverse_direction = False
    assert rc == 0
    snaps, svars = g["states"], g["svars"]
    maxes = [snaps[k][3].max() for k in range(len(snaps))]
    state = np.loadtxt(out, delimiter=",")
    sv = np.loadtxt(tmp_path / "run_svars.txt", delimiter=",")
    assert state.shape == (2 * 4 * 24, 44) and sv.shape == (2 * 4 * 20, 40)
    got = state.reshape(2, 4, 24, 44)
    gsv = sv.reshape(2, 4, 20, 40)
    for k in range(2):
        for v in range(4):
            ref, refs = snaps[k][v], svars[k][v]
            print(f"{ic} snapshot {k} var {v}: state rel {rel_l2(got[k, v], ref):.3e} "
                  f"max|ref| {np.max(np.abs(ref)):.3e} max|diff| "
                  f"{np.max(np.abs(got[k, v] - ref)):.3e}; svars rel {rel_l2(gsv[k, v], refs):.3e} "
                  f"max|ref| {np.max(np.abs(refs)):.3e} max|diff| "
                  f"{np.max(np.abs(gsv[k, v] - refs)):.3e}", flush=True)
    # theta' = (rho*theta' + hyt)/(hyr + rho') - hyt/hyr subtracts two values
    # near the full theta (~300 K), so it carries ~1 ulp of 300 K whatever its
    # inputs; where theta' itself is small (the injection inflow, < 0.2 K) that
    # floor exceeds 1e-12 relative, and theta' is held to 4 ulp of the full
    # theta instead (the states it is computed from agree to ~1e-14)
    p = O.make_params(40, 20, ic)   # the driver's defaults: cfl 1.0, max_speed 500
    assert p["dt"] == float(g["dt"])
    hy = O.init_state(p)
    theta_full = np.max(hy["hy_dens_theta_cell"] / hy["hy_dens_cell"])
    for k in range(2):
        for v in range(4):
            ref = snaps[k][v]
            assert rel_l2(got[k, v], ref) <= 1e-12 if np.any(ref) else np.max(np.abs(got[k, v])) < 1e-12
            refs = svars[k][v]
            if not np.any(refs):
                assert np.max(np.abs(gsv[k, v])) < 1e-12
            elif v == 3:
                assert (rel_l2(gsv[k, v], refs) <= 1e-12 or
                        np.max(np.abs(gsv[k, v] - refs)) <= 4 * np.spacing(theta_full))
            else:
                assert rel_l2(gsv[k, v], refs) <= 1e-12
    text = log.read_text()
    nums = lambda pat: [float(x) for x in re.search(pat, text).groups()]
    gm0, ge0 = nums(r"Start: total_mass, total_energy: (\S+), (\S+)")
    gm1, ge1 = nums(r"End: total_mass, total_energy: (\S+), (\S+)")
    m0, e0, m1, e1 = (float(g[k]) for k in ("mass0", "energy0", "mass1", "energy1"))
    assert abs(gm0 - m0) <= 1e-13 * m0 and abs(ge0 - e0) <= 1e-13 * e0
    assert abs(gm1 - m1) <= 1e-12 * m1 and abs(ge1 - e1) <= 1e-12 * e1
    steps = re.findall(r"Step: (\d+), max\(rho\*t\): (\S+)", text)
    assert [int(s) for s, _ in steps] == [int(x) for x in g["steps"]] == [1, 3]
    for (_, v), ref in zip(steps, maxes):
        assert abs(float(v) - ref) <= 1e-12 * abs(ref)
    assert "Relative change in total_mass, total_energy:" in text

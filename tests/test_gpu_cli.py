"""The command-line driver end to end on the GPU, against the oracle."""
import re

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_cli_outputs_and_log_match_oracle(tmp_path):
    """collision 40x20, 5 steps, output every 2 steps: the state and
    solution-variable files hold the snapshots taken before steps 1 and 3
    (after 1 and 3 steps) in the reference's savetxt layout, and the log
    reports the oracle's start/end mass and energy and max(rho*theta)."""
    from pyminiweather_amd.__main__ import main
    out = tmp_path / "run.txt"
    log = tmp_path / "log.txt"
    rc = main(["--nx", "40", "--nz", "20", "--nsteps", "5", "--output-freq", "2",
               "--ic-type", "collision", "--app-filename", str(out), "--app-log-file", str(log)])
    assert rc == 0
    p = O.make_params(40, 20, "collision")
    orc = O.OracleSolver(p)
    m0, e0 = orc.stats()
    snaps, svars, maxes = [], [], []
    for istep in range(5):
        if (istep + 1) % 2 == 0:
            snaps.append(orc.f["state"].copy())
            svars.append(O.solution_variables(p, orc.f, orc.f["state"]))
            maxes.append(orc.f["state"][3].max())
        orc.timestep()
    m1, e1 = orc.stats()
    state = np.loadtxt(out, delimiter=",")
    sv = np.loadtxt(tmp_path / "run_svars.txt", delimiter=",")
    assert state.shape == (2 * 4 * 24, 44) and sv.shape == (2 * 4 * 20, 40)
    got = state.reshape(2, 4, 24, 44)
    gsv = sv.reshape(2, 4, 20, 40)
    for k in range(2):
        for v in range(4):
            ref = snaps[k][v]
            assert rel_l2(got[k, v], ref) <= 1e-12 if np.any(ref) else np.max(np.abs(got[k, v])) < 1e-12
            refs = svars[k][v]
            assert rel_l2(gsv[k, v], refs) <= 1e-12 if np.any(refs) else np.max(np.abs(gsv[k, v])) < 1e-12
    text = log.read_text()
    nums = lambda pat: [float(x) for x in re.search(pat, text).groups()]
    gm0, ge0 = nums(r"Start: total_mass, total_energy: (\S+), (\S+)")
    gm1, ge1 = nums(r"End: total_mass, total_energy: (\S+), (\S+)")
    assert abs(gm0 - m0) <= 1e-13 * m0 and abs(ge0 - e0) <= 1e-13 * e0
    assert abs(gm1 - m1) <= 1e-12 * m1 and abs(ge1 - e1) <= 1e-12 * e1
    steps = re.findall(r"Step: (\d+), max\(rho\*t\): (\S+)", text)
    assert [int(s) for s, _ in steps] == [1, 3]
    for (_, v), ref in zip(steps, maxes):
        assert abs(float(v) - ref) <= 1e-12 * abs(ref)
    assert "Relative change in total_mass, total_energy:" in text

"""The oracle against the golden vectors written by the reference itself
(tests/golden/make_golden.py). Everything is compared bit for bit: the oracle
keeps the reference's operation order and runs on the same NumPy/glibc."""
import numpy as np
import pytest

import oracle as O
from conftest import load_golden

HYDRO = ("hy_dens_cell", "hy_dens_theta_cell", "hy_dens_int", "hy_dens_theta_int",
         "hy_pressure_int")


@pytest.mark.parametrize("ic,nx,nz", [("collision", 40, 20), ("gravity", 40, 20),
                                      ("density-current", 40, 20), ("injection", 40, 20),
                                      ("thermal", 20, 10)])
def test_init_matches_reference(ic, nx, nz):
    g = load_golden(f"init_{ic}_{nx}x{nz}")
    p = O.make_params(nx, nz, ic, cfl=1.5, max_speed=450.0)
    assert p["dt"] == g["dt"]
    f = O.init_state(p)
    np.testing.assert_array_equal(f["state"], g["state"])
    for k in HYDRO:
        np.testing.assert_array_equal(f[k], g[k])
    m, e = O.stats(p, f, f["state"])
    assert m == g["mass"] and e == g["energy"]


@pytest.mark.parametrize("ic", ["thermal", "gravity"])
@pytest.mark.parametrize("d", ["x", "z"])
@pytest.mark.parametrize("method", ["slice", "scipy"])
def test_discrete_step_matches_reference(ic, d, method):
    g = load_golden(f"discrete_step_{ic}_12x7")
    p = O.make_params(12, 7, ic, dt=float(g["dt_full"]))
    f = O.init_state(p)
    for k in HYDRO:
        np.testing.assert_array_equal(f[k], g[k])
    S = O.OracleSolver(p, f, method=method)
    a, b, c = g["s_init"].copy(), g["s_forc"].copy(), g["s_out"].copy()
    S.stage(a, b, c, float(g["dt_stage"]), d)
    np.testing.assert_array_equal(a, g[f"{d}_init"])
    np.testing.assert_array_equal(b, g[f"{d}_forc"])      # halos written
    np.testing.assert_array_equal(c, g[f"{d}_out"])
    np.testing.assert_array_equal(f["tend"], g[f"{d}_tend"])
    fl = (slice(None), slice(0, 7), slice(0, 13)) if d == "x" else \
        (slice(None), slice(0, 8), slice(0, 12))          # region this direction writes
    np.testing.assert_array_equal(f["flux"][fl], g[f"{d}_flux"][fl])
    a, b = g["s_init"].copy(), g["s_forc"].copy()
    S.stage(a, b, b, float(g["dt_full"]) / 2, d)           # aliased stage 2
    np.testing.assert_array_equal(a, g[f"{d}_alias_init"])
    np.testing.assert_array_equal(b, g[f"{d}_alias_out"])


TRAJ = [("collision", 40, 20, [1, 2, 25]), ("gravity", 40, 20, [1, 10]),
        ("density-current", 40, 20, [1, 10]), ("injection", 40, 20, [1, 10]),
        ("thermal", 100, 50, [1, 2, 10, 1500])]


@pytest.mark.parametrize("ic,nx,nz,steps", TRAJ)
def test_evolve_matches_reference(ic, nx, nz, steps):
    g = load_golden(f"evolve_{ic}_{nx}x{nz}")
    p = O.make_params(nx, nz, ic, cfl=1.5, max_speed=450.0)
    S = O.OracleSolver(p, method="slice")
    np.testing.assert_array_equal(S.f["state"], g["state0"])
    m0, e0 = S.stats()
    assert m0 == g["mass0"] and e0 == g["energy0"]
    done = 0
    for st in steps:
        S.run(st - done)
        done = st
        np.testing.assert_array_equal(S.f["state"], g[f"state_{st}"])
        if f"state_tmp_{st}" in g:
            np.testing.assert_array_equal(S.f["state_tmp"], g[f"state_tmp_{st}"])
        m, e = S.stats()
        assert m == g[f"mass_{st}"] and e == g[f"energy_{st}"]
        # compute_solution_variables (post/stats.py:38-69)
        np.testing.assert_array_equal(O.solution_variables(p, S.f, S.f["state"]),
                                      g[f"svars_{st}"])
    assert S.reverse_direction == bool(g["reverse_direction_end"])


@pytest.mark.parametrize("ic", ["collision", "injection", "density-current"])
def test_driver_output_loop_matches_reference(ic):
    """The driver's output loop (pyminiweather/__main__.py:205-247): the
    states and solution variables written before every output_freq-th step
    after the warm-ups, and the start/end stats, as the reference's own
    functions produced them."""
    g = load_golden(f"cli_{ic}_40x20")
    p = O.make_params(40, 20, ic)   # the driver's defaults: cfl 1.0, max_speed 500
    assert p["dt"] == float(g["dt"])
    S = O.OracleSolver(p)
    m0, e0 = S.stats()
    assert m0 == g["mass0"] and e0 == g["energy0"]
    S.run(int(g["nwarmups"]))
    k = 0
    freq = int(g["output_freq"])
    for istep in range(int(g["nsteps"])):
        if (istep + 1) % freq == 0:
            assert g["steps"][k] == istep
            np.testing.assert_array_equal(S.f["state"], g["states"][k])
            np.testing.assert_array_equal(O.solution_variables(p, S.f, S.f["state"]),
                                          g["svars"][k])
            k += 1
        S.timestep()
    assert k == len(g["steps"])
    m1, e1 = S.stats()
    assert m1 == g["mass1"] and e1 == g["energy1"]


def test_scipy_variant_matches_slice_variant_over_steps():
    """The timed CPU baseline (SciPy convolve) and the parity oracle (slicing)
    are the same arithmetic."""
    p = O.make_params(30, 16, "collision", cfl=1.5, max_speed=450.0)
    A = O.OracleSolver(p, method="scipy")
    B = O.OracleSolver(p, method="slice")
    A.run(3)
    B.run(3)
    np.testing.assert_array_equal(A.f["state"], B.f["state"])
    np.testing.assert_array_equal(A.f["state_tmp"], B.f["state_tmp"])

"""Kernel-shape autotuners and the per-launch timing path they use."""
import ctypes

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


def _fields(p):
    from pyminiweather_amd.data import initialize_fields
    from pyminiweather_amd.ics import init
    from pyminiweather_amd.mesh import MeshData
    f = initialize_fields(p)
    init(f, p, MeshData(p))
    return f


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_evolve_ex_timing_leaves_the_same_state():
    """pmw_evolve_ex with per-launch timing (kernel_ms) gives the untimed
    call's state bit for bit, and positive x and z times."""
    import pyminiweather_amd.solve as S
    from pyminiweather_amd import _lib
    from pyminiweather_amd._lib import ptr
    p = O.make_params(120, 60, "collision", cfl=1.5, max_speed=450.0)
    fa, fb = _fields(p), _fields(p)
    out = []
    for f, timed in ((fa, False), (fb, True)):
        kms = (ctypes.c_double * 2)() if timed else None
        flag = ctypes.c_int(0)
        tune = (ctypes.c_int * 3)(*S.DEFAULT_TUNING)
        _lib.call("pmw_evolve_ex", ptr(f.state), ptr(f.state_tmp), ptr(f.state_work),
                  ptr(f.hy_dens_cell), ptr(f.hy_dens_theta_cell), ptr(f.hy_dens_int),
                  ptr(f.hy_dens_theta_int), ptr(f.hy_pressure_int), None, 0, None, f.nx, f.nz,
                  p["dx"], p["dz"], p["dt"], S.hv_coeff(p["dx"], p["dt"]),
                  S.hv_coeff(p["dz"], p["dt"]), ctypes.byref(flag), 4, tune,
                  _lib.stream_of(), kms)
        out.append(_host(f.state))
        if timed:
            assert kms[0] > 0.0 and kms[1] > 0.0
    np.testing.assert_array_equal(out[0], out[1])


def test_tune_grid_registers_a_candidate_and_leaves_the_state():
    import pyminiweather_amd.solve as S
    from pyminiweather_amd.autotune import X_BLOCKS, Z_SEGMENTS, tune_grid
    p = O.make_params(96, 44, "thermal", cfl=1.5, max_speed=450.0)
    f = _fields(p)
    before = _host(f.state).copy()
    best, table = tune_grid(p, f)
    try:
        assert best[0] in (0, 1) and best[1] in X_BLOCKS and best[2] in Z_SEGMENTS
        assert S.tuning_for(96, 44) == best
        assert best in table and all(x > 0 and z > 0 for x, z in table.values())
        np.testing.assert_array_equal(before, _host(f.state))
    finally:
        S._tuning_table.pop((96, 44), None)


def test_tune_ensemble_picks_a_fused_shape_that_keeps_members_exact():
    """After tune_ensemble the ensemble still equals the single solves of its
    members bit for bit (the tuning runs advance the ensemble, so the
    comparison starts from its state after tuning)."""
    import pyminiweather_amd.solve as S
    from pyminiweather_amd.autotune import X_BLOCKS, Z_SEGMENTS, tune_ensemble
    from pyminiweather_amd.ensemble import Ensemble, member_params
    params = [member_params(40, 20), member_params(64, 30, "collision"),
              member_params(200, 100)]
    ens = Ensemble(params, [None, None, None])
    best, table = tune_ensemble(ens)
    assert best[0] == 0 and best[1] in X_BLOCKS and best[2] in Z_SEGMENTS
    assert ens.tuning == best and all(x > 0 and z > 0 for x, z in table.values())
    starts = [(f.state.clone(), f.state_tmp.clone()) for f in ens.fields]
    flag = ens.reverse_direction
    ens.evolve(3)
    saved = S.reverse_direction
    try:
        for (s0, t0), p, f in zip(starts, params, ens.fields):
            g = _fields(p)
            g.state.copy_(s0)
            g.state_tmp.copy_(t0)
            S.set_tuning(p["nx"], p["nz"], best)
            S.reverse_direction = flag
            S.run(p, g, None, 3)
            S._tuning_table.pop((p["nx"], p["nz"]), None)
            np.testing.assert_array_equal(_host(g.state), _host(f.state))
    finally:
        S.reverse_direction = saved
        for p in params:
            S._tuning_table.pop((p["nx"], p["nz"]), None)

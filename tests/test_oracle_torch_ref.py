"""The FP64 torch restatement (autograd reference of the adjoint) against
the NumPy oracle, on the CPU."""
import numpy as np
import pytest
import torch

import oracle as O
from oracle import torch_ref as R
from conftest import rel_l2


@pytest.mark.parametrize("ic,nx,nz,steps", [("thermal", 24, 12, 6), ("collision", 20, 10, 5),
                                            ("gravity", 16, 8, 4), ("density-current", 14, 9, 3)])
def test_torch_ref_trajectory_matches_oracle(ic, nx, nz, steps):
    p = O.make_params(nx, nz, ic, cfl=1.5, max_speed=450.0)
    orc = O.OracleSolver(p)
    f = {k: torch.from_numpy(orc.f[k].copy()) for k in
         ("hy_dens_cell", "hy_dens_theta_cell", "hy_dens_int", "hy_dens_theta_int",
          "hy_pressure_int")}
    src = None
    if ic == "gravity":
        # the w source the oracle adds every stage, isolated from a rest state
        rest = O.OracleSolver(p)
        zero = np.zeros_like(rest.f["state"])
        out = np.zeros_like(zero)
        rest.stage(zero, zero.copy(), out, 1.0, "z")
        grav = O.OracleSolver(dict(p, ic_type="thermal"), fields=rest.f)
        out2 = np.zeros_like(zero)
        grav.stage(zero, zero.copy(), out2, 1.0, "z")
        src = torch.from_numpy((out - out2)[2, 2:-2, 2:-2].copy())
    u0 = torch.from_numpy(orc.f["state"][:, 2:-2, 2:-2].copy())
    ut = R.evolve(u0, p, f, steps, src_w=src)
    orc.run(steps)
    ref = orc.f["state"][:, 2:-2, 2:-2]
    for v in range(4):
        if np.any(ref[v]):
            assert rel_l2(ut[v].numpy(), ref[v]) <= 1e-11, (v, rel_l2(ut[v].numpy(), ref[v]))
        else:
            assert np.max(np.abs(ut[v].numpy())) < 1e-12


def test_torch_ref_single_stage_tendencies_match_oracle():
    """tend_x / tend_z on a random state equal the oracle's per-stage
    tendencies (flux, hyperviscosity, walls, gravity)."""
    p = O.make_params(12, 7, "collision", cfl=1.5, max_speed=450.0)
    orc = O.OracleSolver(p)
    rng = np.random.default_rng(3)
    s = orc.f["state"].copy()
    s[:, 2:-2, 2:-2] += rng.normal(size=s[:, 2:-2, 2:-2].shape) * np.array(
        [1e-3, 2.0, 2.0, 0.5])[:, None, None]
    f = {k: torch.from_numpy(orc.f[k].copy()) for k in
         ("hy_dens_cell", "hy_dens_theta_cell", "hy_dens_int", "hy_dens_theta_int",
          "hy_pressure_int")}
    hvx = -0.05 * p["dx"] / (16 * p["dt"])
    hvz = -0.05 * p["dz"] / (16 * p["dt"])
    for d, tfun, dd, hv in (("x", R.tend_x, p["dx"], hvx), ("z", R.tend_z, p["dz"], hvz)):
        out = np.zeros_like(s)
        init = np.zeros_like(s)
        orc.stage(init, s.copy(), out, 1.0, d)
        t = tfun(torch.from_numpy(s[:, 2:-2, 2:-2].copy()), f, dd, hv).numpy()
        for v in range(4):
            assert rel_l2(t[v], out[v, 2:-2, 2:-2]) <= 1e-13

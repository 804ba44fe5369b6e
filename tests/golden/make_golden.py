"""Write the golden fixtures under tests/golden/ by running the reference.

This script is the only place that imports the reference PyMiniWeather, and it
only ever runs in the development container (the reference does not travel):

    python tests/golden/make_golden.py /root/reference

It pins what no closed form or invariant can pin:

* ``init`` (pyminiweather/ics/initial.py:11-105) for every ``--ic-type``; its
  (rho*theta)' formula cancels catastrophically (initial.py:73-77) and its
  quadrature sum order is NumPy's.
* ``discrete_step`` (pyminiweather/solve/step.py:21-82) on a seeded, perturbed
  state with three distinct buffers and with forcing == out (the aliased RK
  stage 2, step.py:122-131); hyperviscosity uses ``params["dt"]``, not the stage
  dt (solve/interpolate.py:99-101, 147-149).
* ``evolve`` (pyminiweather/solve/step.py:85-143) trajectories, Z-first on the
  first call with the module-global toggle (step.py:18, 103-109, 143).
* ``compute_stats`` (pyminiweather/post/stats.py:8-35) on those states.
* ``compute_solution_variables`` (pyminiweather/post/stats.py:38-69) on
  every evolve checkpoint, and the command-line driver's output loop
  (pyminiweather/__main__.py:205-247: warm-ups, then the state and the
  solution variables before every ``output_freq``-th step) for three short
  runs (``cli_*``).

Output: ``tests/golden/*.npz`` (compressed float64 arrays + scalars). The
thermal 100x50 ``init`` output is ``state0`` of its evolve trajectory.
"""
import os
import sys
import time

import numpy as np


def _params(ic_type, nx, nz, cfl=1.5, max_speed=450.0, dt=None):
    # pyminiweather/__main__.py:160-195: dx, dz, then dt from the CFL condition
    p = {
        "nx": nx, "nz": nz, "xlen": 2.0e4, "zlen": 1.0e4, "hs": 2, "s": 4,
        "ic_type": ic_type, "cfl": cfl, "max_speed": max_speed,
        "nsteps": 0, "nwarmups": 0, "output_freq": -1, "verbose": False,
    }
    p["dx"] = p["xlen"] / p["nx"]
    p["dz"] = p["zlen"] / p["nz"]
    p["dt"] = dt if dt is not None else np.minimum(p["dx"], p["dz"]) * cfl / max_speed
    return p


def main(ref_root):
    for var in ("LEGATE_MAX_DIM", "LEGATE_MAX_FIELDS"):
        os.environ.pop(var, None)
    sys.path.insert(0, ref_root)
    import pyminiweather.solve.step as step_mod
    from pyminiweather.data import initialize_fields
    from pyminiweather.ics import Directions, init
    from pyminiweather.mesh import MeshData
    from pyminiweather.post import compute_solution_variables, compute_stats
    from pyminiweather.solve import evolve

    out_dir = os.path.dirname(os.path.abspath(__file__))

    def setup(ic_type, nx, nz, **kw):
        params = _params(ic_type, nx, nz, **kw)
        fields = initialize_fields(params)
        mesh = MeshData(params)
        init(fields, params, mesh)
        return params, fields, mesh

    def hydro(fields):
        return dict(
            hy_dens_cell=fields.hy_dens_cell.copy(),
            hy_dens_theta_cell=fields.hy_dens_theta_cell.copy(),
            hy_dens_int=fields.hy_dens_int.copy(),
            hy_dens_theta_int=fields.hy_dens_theta_int.copy(),
            hy_pressure_int=fields.hy_pressure_int.copy(),
        )

    def save(name, **arrays):
        path = os.path.join(out_dir, name + ".npz")
        np.savez_compressed(path, **arrays)
        print(f"wrote {path} ({os.path.getsize(path)} B)", flush=True)

    # ---- init for every IC type (small grids) -------------------------------
    for ic, nx, nz in [("collision", 40, 20),
                       ("gravity", 40, 20), ("density-current", 40, 20),
                       ("injection", 40, 20), ("thermal", 20, 10)]:
        params, fields, _ = setup(ic, nx, nz)
        mass, energy = compute_stats(params, fields)
        assert np.array_equal(fields.state, fields.state_tmp)
        save(f"init_{ic}_{nx}x{nz}", state=fields.state.copy(), mass=float(mass),
             energy=float(energy), dt=float(params["dt"]), **hydro(fields))

    # ---- discrete_step on a seeded perturbed state --------------------------
    nx, nz = 12, 7
    for ic in ("thermal", "gravity"):
        params, fields, mesh = setup(ic, nx, nz, dt=0.37)
        rng = np.random.default_rng(7)
        base = fields.state.copy()
        scale = np.array([1e-3, 2.0, 2.0, 0.5])[:, None, None]
        s_init = base + scale * rng.standard_normal(base.shape)
        s_forc = base + scale * rng.standard_normal(base.shape)
        s_out = base + scale * rng.standard_normal(base.shape)
        rec = dict(s_init=s_init.copy(), s_forc=s_forc.copy(), s_out=s_out.copy(),
                   dt_full=0.37, dt_stage=0.37 / 3, **hydro(fields))
        for dname, d in (("x", Directions.X), ("z", Directions.Z)):
            a, b, c = s_init.copy(), s_forc.copy(), s_out.copy()
            step_mod.discrete_step(params, fields, mesh, a, b, c, 0.37 / 3, d)
            rec[f"{dname}_init"] = a
            rec[f"{dname}_forc"] = b
            rec[f"{dname}_out"] = c
            rec[f"{dname}_tend"] = fields.tend.copy()
            rec[f"{dname}_flux"] = fields.flux.copy()
            # aliased stage: forcing and out are the same buffer
            a, b = s_init.copy(), s_forc.copy()
            step_mod.discrete_step(params, fields, mesh, a, b, b, 0.37 / 2, d)
            rec[f"{dname}_alias_init"] = a
            rec[f"{dname}_alias_out"] = b
        save(f"discrete_step_{ic}_{nx}x{nz}", **rec)

    # ---- evolve trajectories ------------------------------------------------
    def trajectory(ic, nx, nz, checkpoints, keep_tmp):
        step_mod.reverse_direction = False
        params, fields, mesh = setup(ic, nx, nz)
        m0, e0 = compute_stats(params, fields)
        rec = dict(state0=fields.state.copy(), dt=float(params["dt"]),
                   mass0=float(m0), energy0=float(e0), **hydro(fields))
        t0 = time.perf_counter()
        nmax = max(checkpoints)
        for istep in range(1, nmax + 1):
            evolve(params, fields, mesh, dt=params["dt"])
            if istep in checkpoints:
                rec[f"state_{istep}"] = fields.state.copy()
                rec[f"svars_{istep}"] = np.array(compute_solution_variables(params, fields))
                if istep in keep_tmp:
                    rec[f"state_tmp_{istep}"] = fields.state_tmp.copy()
                m, e = compute_stats(params, fields)
                rec[f"mass_{istep}"] = float(m)
                rec[f"energy_{istep}"] = float(e)
        rec["reverse_direction_end"] = bool(step_mod.reverse_direction)
        print(f"{ic} {nx}x{nz}: {nmax} steps in {time.perf_counter() - t0:.1f} s",
              flush=True)
        return rec

    save("evolve_collision_40x20",
         **trajectory("collision", 40, 20, {1, 2, 25}, {1, 2, 25}))
    save("evolve_gravity_40x20", **trajectory("gravity", 40, 20, {1, 10}, {1}))
    save("evolve_density-current_40x20",
         **trajectory("density-current", 40, 20, {1, 10}, {1}))
    save("evolve_injection_40x20",
         **trajectory("injection", 40, 20, {1, 10}, {1, 10}))
    save("evolve_thermal_100x50",
         **trajectory("thermal", 100, 50, {1, 2, 10, 1500}, {1}))

    # ---- the driver's output loop (__main__.py:205-247) ---------------------
    def cli_run(ic, nx, nz, nsteps, output_freq, nwarmups):
        # the driver's defaults: --cfl 1.0 --max-speed 500 (__main__.py:109-122)
        step_mod.reverse_direction = False
        params, fields, mesh = setup(ic, nx, nz, cfl=1.0, max_speed=500.0)
        m0, e0 = compute_stats(params, fields)
        for _ in range(nwarmups):
            evolve(params, fields, mesh, dt=params["dt"])
        states, svars, steps = [], [], []
        for istep in range(nsteps):
            if output_freq > 0 and (istep + 1) % output_freq == 0:
                states.append(fields.state.copy())
                svars.append(np.array(compute_solution_variables(params, fields)))
                steps.append(istep)
            evolve(params, fields, mesh, dt=params["dt"])
        m1, e1 = compute_stats(params, fields)
        return dict(states=np.stack(states), svars=np.stack(svars), steps=np.array(steps),
                    mass0=float(m0), energy0=float(e0), mass1=float(m1), energy1=float(e1),
                    nsteps=nsteps, output_freq=output_freq, nwarmups=nwarmups,
                    dt=float(params["dt"]))

    for ic, warmups in (("collision", 0), ("injection", 0), ("density-current", 2)):
        save(f"cli_{ic}_40x20", **cli_run(ic, 40, 20, 5, 2, warmups))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

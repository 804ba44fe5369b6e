"""Central finite differences of the theta' misfit computed entirely by the
oracle (NumPy, the reference's own arithmetic) on a grid small enough to run
here, to show where the FP64 floor of such a quotient lies.

    python tests/golden/make_fd_100x50.py        # ~1 min, 4 processes

Problem (the 400x200 adjoint workload shrunk to 100x50, same 400 steps):
thermal IC (ics/initial_conditions.py:166-199), cfl 1.5, max_speed 450;
theta0 = the IC's (rho*theta)' divided by hy_dens_cell per row; the control
maps theta' to the initial (rho*theta)' = hy_dens_cell[k] theta' with the
other perturbations zero; target = the oracle's final (rho*theta)' from
theta0 + N(0, 0.1 K) noise (numpy default_rng(1)); L = 1/2 ||(rho*theta)'(T)
- target||^2 over the interior after 400 evolve steps
(pyminiweather/solve/step.py:85-143).

Written to ``fd_thermal_100x50.npz``: ``theta0``, ``target``, ``loss``,
``cells`` (16 x (k, i), default_rng(7)), ``eps`` (1e-6, 1e-5, 1e-4, 1e-3),
``fd`` (len(eps) x 16 central quotients), ``jitter`` (L(theta0 + 1e-13 e_c)
- L(theta0) for the first 8 cells: the loss's own round-off noise).
"""
import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

NX, NZ, STEPS = 100, 50, 400
EPS = (1e-6, 1e-5, 1e-4, 1e-3)


def _params():
    import oracle as O
    return O.make_params(NX, NZ, "thermal", cfl=1.5, max_speed=450.0)


def final_rhot(theta):
    """Oracle forward from the theta' control; returns the final interior
    (rho*theta)'."""
    import oracle as O
    p = _params()
    S = O.OracleSolver(p)
    w = S.f["hy_dens_cell"][2:NZ + 2, None]
    for key in ("state", "state_tmp"):
        S.f[key][...] = 0.0
    S.f["state"][O.RHOT, 2:-2, 2:-2] = w * theta
    S.run(STEPS)
    return S.f["state"][O.RHOT, 2:-2, 2:-2].copy()


def _loss(args):
    theta, target = args
    r = final_rhot(theta) - target
    return 0.5 * float(np.sum(r * r))


def main():
    import oracle as O
    p = _params()
    f0 = O.init_state(p)
    theta0 = f0["state"][O.RHOT, 2:-2, 2:-2] / f0["hy_dens_cell"][2:NZ + 2, None]
    noise = np.random.default_rng(1).normal(0.0, 0.1, (NZ, NX))
    target = final_rhot(theta0 + noise)
    rng = np.random.default_rng(7)
    cells = np.array([(int(rng.integers(NZ)), int(rng.integers(NX))) for _ in range(16)])
    jobs = [(theta0, target)]
    for eps in EPS:
        for k, i in cells:
            for sgn in (1.0, -1.0):
                th = theta0.copy()
                th[k, i] += sgn * eps
                jobs.append((th, target))
    for k, i in cells[:8]:
        th = theta0.copy()
        th[k, i] += 1e-13
        jobs.append((th, target))
    with mp.get_context("spawn").Pool(4) as pool:
        losses = pool.map(_loss, jobs)
    base = losses[0]
    fd = np.zeros((len(EPS), len(cells)))
    j = 1
    for a, eps in enumerate(EPS):
        for c in range(len(cells)):
            fd[a, c] = (losses[j] - losses[j + 1]) / (2 * eps)
            j += 2
    jitter = np.array(losses[j:]) - base
    out = os.path.join(HERE, "fd_thermal_100x50.npz")
    np.savez_compressed(out, theta0=theta0, target=target, loss=base, cells=cells,
                        eps=np.array(EPS), fd=fd, jitter=jitter)
    print("loss", base, "sigma_L", float(np.sqrt(np.mean(jitter ** 2))))
    for a, eps in enumerate(EPS):
        print(f"eps {eps:g}: fd[:4] {fd[a, :4]}")
    print(f"wrote {out} ({os.path.getsize(out)} B)", flush=True)


if __name__ == "__main__":
    main()

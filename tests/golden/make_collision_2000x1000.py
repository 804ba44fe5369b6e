"""Reduced fixture of the 2000x1000 colliding-thermals workload after 500 steps.

The trajectory is far too long for the reference itself to be run here in
reasonable time (about 13 s per step through scipy.signal.convolve), so it is
produced by the oracle's slicing variant, which reproduces the reference bit
for bit: it matches every reference-written golden vector exactly
(tests/test_oracle_golden.py), including the 1500-step thermal 100x50
trajectory, and the slicing and SciPy variants are bit-identical.

    python tests/golden/make_collision_2000x1000.py [KEEP_DIR]   # ~30 min, 2 processes

(KEEP_DIR, if given, keeps the full states of both runs as .npy files.)

Reference path: pyminiweather/solve/step.py:85-143 (evolve), driven as by
pyminiweather/__main__.py:217-237 (Z-first toggle, nsteps calls); the IC is
ics/initial_conditions.py:96-134 (collision) with cfl 1.5 and max_speed 450.

Two runs go side by side in separate processes:

* the clean oracle;
* the same oracle with 1% of its flux pow results moved by one ulp (seeded),
  which measures how far the reference's own arithmetic moves under
  last-bit pow differences: the floor below which an implementation with a
  different (even correctly rounded) pow cannot be held.

Written to ``collision_2000x1000_500.npz`` (about 1 MB):

* ``rows``, ``cols``: the sampled interior rows and columns (every 16th);
* ``sample_<n>``: state[:, 2 + rows, 2 + cols] of the clean run at step n;
* ``norm_<n>``: per-variable L2 norm of the clean interior at step n;
* ``mass_<n>``, ``energy_<n>``: post/stats.py:8-35 on the clean state;
* ``noise_rel_<n>``: per-variable relative L2 (whole interior) of the noisy
  run against the clean one at step n;
* ``dt``, ``steps``.
"""
import multiprocessing as mp
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

NX, NZ = 2000, 1000
CHECKPOINTS = (100, 250, 500)
STRIDE = 16
NOISE_RATE = 0.01


def _run(noisy, out_dir):
    import oracle as O
    p = O.make_params(NX, NZ, "collision", cfl=1.5, max_speed=450.0)
    if noisy:
        rng = np.random.default_rng(11)
        power = np.power

        def noisy_power(a, b, *args, **kw):
            r = power(a, b, *args, **kw)
            if isinstance(r, np.ndarray) and r.size > 100:
                flip = rng.random(r.shape) < NOISE_RATE
                up = rng.random(r.shape) < 0.5
                r = np.where(flip, np.nextafter(r, np.where(up, np.inf, -np.inf)), r)
            return r
        np.power = noisy_power
    S = O.OracleSolver(p)
    tag = "noisy" if noisy else "clean"
    t0 = time.perf_counter()
    done = 0
    for n in CHECKPOINTS:
        S.run(n - done)
        done = n
        np.save(os.path.join(out_dir, f"{tag}_{n}.npy"), S.f["state"])
        if not noisy:
            m, e = S.stats()
            np.save(os.path.join(out_dir, f"stats_{n}.npy"), np.array([m, e]))
        print(f"{tag}: step {n} after {time.perf_counter() - t0:.0f} s", flush=True)


def main(keep_dir=None):
    out = os.path.join(HERE, "collision_2000x1000_500.npz")
    with tempfile.TemporaryDirectory() as tmp_dir:
        tmp = keep_dir or tmp_dir
        os.makedirs(tmp, exist_ok=True)
        procs = [mp.Process(target=_run, args=(noisy, tmp)) for noisy in (False, True)]
        for pr in procs:
            pr.start()
        for pr in procs:
            pr.join()
            if pr.exitcode != 0:
                raise SystemExit(f"oracle run failed with exit code {pr.exitcode}")
        import oracle as O
        p = O.make_params(NX, NZ, "collision", cfl=1.5, max_speed=450.0)
        rows = np.arange(0, NZ, STRIDE)
        cols = np.arange(0, NX, STRIDE)
        rec = {"rows": rows, "cols": cols, "dt": float(p["dt"]),
               "steps": np.array(CHECKPOINTS)}
        for n in CHECKPOINTS:
            clean = np.load(os.path.join(tmp, f"clean_{n}.npy"))
            noisy = np.load(os.path.join(tmp, f"noisy_{n}.npy"))
            ci = clean[:, 2:-2, 2:-2]
            ni = noisy[:, 2:-2, 2:-2]
            rec[f"sample_{n}"] = ci[:, rows][:, :, cols].copy()
            rec[f"norm_{n}"] = np.array([np.linalg.norm(ci[v]) for v in range(4)])
            rec[f"noise_rel_{n}"] = np.array(
                [np.linalg.norm(ni[v] - ci[v]) / np.linalg.norm(ci[v]) for v in range(4)])
            m, e = np.load(os.path.join(tmp, f"stats_{n}.npy"))
            rec[f"mass_{n}"] = float(m)
            rec[f"energy_{n}"] = float(e)
            print(f"step {n}: noise rel L2 per var",
                  " ".join(f"{x:.2e}" for x in rec[f"noise_rel_{n}"]), flush=True)
    np.savez_compressed(out, **rec)
    print(f"wrote {out} ({os.path.getsize(out)} B)", flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else None)

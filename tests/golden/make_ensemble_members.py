"""Reduced fixtures of one member of each grid size of the seeded 256-member
mixed-size thermal ensemble, after 300 steps.

    python tests/golden/make_ensemble_members.py      # ~1.5 min, 6 processes

The members are the first of each size in the seed-0 draw
(pyminiweather_amd.ensemble.mixed_thermal_members: per member a grid from
200x100 / 400x200 / 800x400, then amplitude ~U(1,5) K, x0 ~U(6,14) km,
z0 ~U(1.5,3) km, radius ~U(1.5,2.5) km). Each is advanced 300 steps from its
thermal IC (ics/initial_conditions.py:166-199 with the drawn bubble, its own
CFL dt) by the oracle's slicing variant, which reproduces the reference's
evolve (pyminiweather/solve/step.py:85-143) bit for bit on every
reference-written golden vector. Beside each clean run, the same run with 1%
of the flux pow results moved by one ulp measures how far the reference's
own arithmetic moves under last-bit pow differences.

Written to ``ensemble_members_300.npz``: for member size tag ``<nx>x<nz>``,
``<tag>_index`` (position in the 256-member draw), ``<tag>_bubble``,
``<tag>_stride``, ``<tag>_sample`` (interior state every stride-th row and
column), ``<tag>_norm`` (per-variable interior L2 norms),
``<tag>_noise_rel`` (per-variable relative L2 of the noisy run against the
clean one, whole interior), ``<tag>_mass``, ``<tag>_energy``.
"""
import multiprocessing as mp
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

STEPS = 300
STRIDES = {(200, 100): 2, (400, 200): 4, (800, 400): 8}
NOISE_RATE = 0.01


def first_members():
    from pyminiweather_amd.ensemble import mixed_thermal_members
    params, bubbles = mixed_thermal_members(256, seed=0)
    out = {}
    for q, (p, b) in enumerate(zip(params, bubbles)):
        out.setdefault((p["nx"], p["nz"]), (q, p, b))
    return out


def _run(args):
    p, bubble, noisy, path = args
    import oracle as O
    if noisy:
        rng = np.random.default_rng(13)
        power = np.power

        def noisy_power(a, b, *a2, **kw):
            r = power(a, b, *a2, **kw)
            if isinstance(r, np.ndarray) and r.size > 100:
                flip = rng.random(r.shape) < NOISE_RATE
                up = rng.random(r.shape) < 0.5
                r = np.where(flip, np.nextafter(r, np.where(up, np.inf, -np.inf)), r)
            return r
        np.power = noisy_power
    S = O.OracleSolver(p, bubble=bubble)
    S.run(STEPS)
    np.save(path, S.f["state"])
    if not noisy:
        m, e = S.stats()
        np.save(path + ".stats.npy", np.array([m, e]))


def main():
    members = first_members()
    rec = {}
    with tempfile.TemporaryDirectory() as tmp:
        jobs = []
        for (nx, nz), (q, p, b) in members.items():
            for noisy in (False, True):
                jobs.append((p, b, noisy, os.path.join(tmp, f"{nx}x{nz}_{int(noisy)}.npy")))
        ctx = mp.get_context("spawn")
        with ctx.Pool(len(jobs)) as pool:
            pool.map(_run, jobs)
        for (nx, nz), (q, p, b) in sorted(members.items()):
            tag = f"{nx}x{nz}"
            clean = np.load(os.path.join(tmp, f"{tag}_0.npy"))[:, 2:-2, 2:-2]
            noisy = np.load(os.path.join(tmp, f"{tag}_1.npy"))[:, 2:-2, 2:-2]
            st = STRIDES[(nx, nz)]
            m, e = np.load(os.path.join(tmp, f"{tag}_0.npy.stats.npy"))
            rec[f"{tag}_index"] = q
            rec[f"{tag}_bubble"] = np.array(b)
            rec[f"{tag}_stride"] = st
            rec[f"{tag}_sample"] = clean[:, ::st, ::st].copy()
            rec[f"{tag}_norm"] = np.array([np.linalg.norm(clean[v]) for v in range(4)])
            rec[f"{tag}_noise_rel"] = np.array(
                [np.linalg.norm(noisy[v] - clean[v]) / np.linalg.norm(clean[v])
                 for v in range(4)])
            rec[f"{tag}_mass"] = float(m)
            rec[f"{tag}_energy"] = float(e)
            print(f"{tag} (member {q}): noise rel L2",
                  " ".join(f"{x:.2e}" for x in rec[f"{tag}_noise_rel"]), flush=True)
    out = os.path.join(HERE, "ensemble_members_300.npz")
    np.savez_compressed(out, **rec)
    print(f"wrote {out} ({os.path.getsize(out)} B)", flush=True)


if __name__ == "__main__":
    main()

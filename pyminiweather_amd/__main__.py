"""Command-line driver: ``python -m pyminiweather_amd`` (mirrors
pyminiweather/__main__.py:28-248).

Same flags, defaults and choices, the same params dict, the same log lines
(start/end total mass and energy, their relative change, the per-output
"Step: i, max(rho*t): ..." line, the TimedCodeBlock timings) and the same
output files and layout. Warm-up steps advance the state and the direction
toggle as in the reference. The steps between two outputs are handed to the
library as one call (bitwise the same as one evolve() per step), and each
snapshot is copied to pinned host memory while the following steps run.
"""
import argparse
import logging
import sys
from pathlib import Path

import numpy as np


def get_logger(filename=None):
    logger = logging.getLogger("pyminiweather")
    for h in list(logger.handlers):  # a second run in one process logs to its own target only
        logger.removeHandler(h)
        h.close()
    handler = logging.StreamHandler() if filename is None else logging.FileHandler(
        filename, mode="w")
    handler.setFormatter(logging.Formatter("%(message)s"))
    logger.addHandler(handler)
    logger.setLevel(logging.INFO)
    return logger


def get_parser():
    p = argparse.ArgumentParser()
    p.add_argument("--nx", type=int, default=200, dest="nx",
                   help="Number of points in x-direction (default: 200)")
    p.add_argument("--nz", type=int, default=100, dest="nz",
                   help="Number of points in z-direction (default: 100)")
    p.add_argument("--xlen", type=float, default=2e4, dest="xlen",
                   help="Length of domain in x-direction (default: 2e4)")
    p.add_argument("--zlen", type=float, default=1e4, dest="zlen",
                   help="Length of domain in Z-direction (default: 1e4)")
    p.add_argument("--nsteps", type=int, default=10, dest="nsteps",
                   help="Number of time steps (default: 10)")
    p.add_argument("--dt", type=float, default=None, dest="dt",
                   help="Time step size (if none, calculated using CFL)")
    p.add_argument("--nwarmups", type=int, default=0, dest="nwarmups",
                   help="Number of wamup time steps (default: 0)")
    p.add_argument("--ic-type", type=str, default="thermal",
                   choices=["thermal", "collision", "gravity", "density-current",
                            "injection"],
                   help="Type of initial condition. The default value is thermal.",
                   dest="ic_type")
    p.add_argument("--hs", type=int, default=2, choices=[2], dest="hs",
                   help="Width of the stencil halo (fixed)")
    p.add_argument("--s", type=int, default=4, choices=[4], dest="s",
                   help="Size of the interpolating kernel (fixed)")
    p.add_argument("--max-speed", type=float, default=500.0, dest="max_speed",
                   help="Assumed maximum speed for the time-step (default: 500)")
    p.add_argument("--cfl", type=float, default=1.00, dest="cfl",
                   help="CFL number for the time-step (default: 1.0)")
    p.add_argument("--verbose", action="store_true", default=False, dest="verbose",
                   help="Print the simulation setup")
    p.add_argument("--output-freq", type=int, default=-1, dest="output_freq",
                   help="Write the solution every output-freq steps (default: -1, off)")
    p.add_argument("--app-filename", type=str, default="PyMiniWeatherData.txt",
                   dest="app_filename",
                   help="Output file (default: PyMiniWeatherData.txt)")
    p.add_argument("--app-log-file", action="store", type=Path, default=None,
                   metavar="FILE", help="Write messages from the application to a file")
    return p


def get_params_from_args(args):
    keys = ("nx", "nz", "xlen", "zlen", "dt", "nsteps", "nwarmups", "ic_type", "hs", "s",
            "max_speed", "cfl", "output_freq", "app_filename", "app_log_file", "verbose")
    return {k: getattr(args, k) for k in keys}


def derived_params(params):
    """dx, dz and the CFL time step exactly as pyminiweather/__main__.py:190-195."""
    params["dx"] = params["xlen"] / params["nx"]
    params["dz"] = params["zlen"] / params["nz"]
    if params["dt"] is None:
        params["dt"] = np.minimum(params["dx"], params["dz"]) * params["cfl"] / \
            params["max_speed"]
    return params


def main(argv=None):
    args, _ = get_parser().parse_known_args(argv)
    params = derived_params(get_params_from_args(args))
    logger = get_logger(args.app_log_file)
    if params["verbose"]:
        for k, v in params.items():
            logger.info(f"{k:25s} {v}")

    from . import solve as S
    from .data import initialize_fields
    from .ics import init
    from .mesh import MeshData
    from .output import SnapshotWriter
    from .post import compute_stats
    from .utils import TimedCodeBlock

    with TimedCodeBlock(label="Elapsed time for initialization"):
        fields = initialize_fields(params)
        mesh = MeshData(params)
        init(fields, params, mesh)

    mass0, energy0 = compute_stats(params, fields)
    logger.info(f"Start: total_mass, total_energy: {mass0}, {energy0}")

    if params["nwarmups"]:
        with TimedCodeBlock(label="Elapsed time for warmups"):
            S.run(params, fields, mesh, params["nwarmups"])

    freq = params["output_freq"]
    nsteps = params["nsteps"]
    writer = SnapshotWriter(params, fields, params["app_filename"]) if freq > 0 else None
    with TimedCodeBlock(label="Elapsed time for timestepping"):
        istep = 0
        while istep < nsteps:
            if writer is not None and (istep + 1) % freq == 0:
                writer.capture()
                logger.info(f"Step: {istep}, max(rho*t): {writer.max_rho_theta()}")
            # steps up to (not including) the next output step j, (j+1) % freq == 0
            nxt = nsteps
            if writer is not None:
                r = (istep + 1) % freq
                nxt = min(nsteps, istep + (freq if r == 0 else freq - r))
            S.run(params, fields, mesh, nxt - istep)
            if writer is not None:
                writer.flush()  # formats the snapshot while those steps run
            istep = nxt

    mass1, energy1 = compute_stats(params, fields)
    logger.info(f"End: total_mass, total_energy: {mass1}, {energy1}")
    logger.info(f"Relative change in total_mass, total_energy: {(mass1 - mass0) / mass0}, "
                f"{(energy1 - energy0) / energy0}")
    return 0


if __name__ == "__main__":
    sys.exit(main())

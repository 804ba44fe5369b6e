"""Kernel-shape autotuner for the fused time step.

For one grid (or one ensemble) it times every candidate shape with HIP
events around each launch (``pmw_evolve_ex`` / ``pmw_ensemble_evolve``
``kernel_ms``) and keeps, independently, the fastest x block size and the
fastest z segment length (the two directions are separate launches), then
checks the fused choice against one launch per RK stage. Single grids are
timed on scratch copies of the state, so tuning does not advance the
solution; ensembles are timed in place (tuning advances them, like warm-up
steps).
"""
import ctypes
import itertools

import torch

from . import _lib
from . import solve as S
from ._lib import ptr, stream_of

# x: threads per row tile, or 1000 + threads for dense tiles (rows end to end)
X_BLOCKS = (256, 448, 512, 1256, 1512)
# z: rows per marching segment, or an LDS tile code columns*1000 + rows
# (8064, 4128: one cell per thread; 8096: one cell per thread with the RK
# init in registers; 8128: two cells per thread; 18128 / 18096: 8128 /
# 8096 as dense tiles, column groups end to end)
Z_SEGMENTS = (8, 16, 24, 32, 48, 64, 8064, 4128, 8096, 8128, 18128, 18096)


def _time_grid(params, fields, tuning, steps, bufs):
    kms = (ctypes.c_double * 2)()
    flag = ctypes.c_int(0)
    tune = (ctypes.c_int * 3)(*tuning)
    bc = 1 if params["ic_type"] == "injection" else 0
    _lib.call("pmw_evolve_ex", ptr(bufs[0]), ptr(bufs[1]), ptr(bufs[2]),
              ptr(fields.hy_dens_cell), ptr(fields.hy_dens_theta_cell),
              ptr(fields.hy_dens_int), ptr(fields.hy_dens_theta_int),
              ptr(fields.hy_pressure_int),
              ptr(fields.src_w) if params["ic_type"] == "gravity" else None, bc,
              ptr(fields.inj_rows) if bc else None, fields.nx, fields.nz,
              float(params["dx"]), float(params["dz"]), float(params["dt"]),
              S.hv_coeff(params["dx"], params["dt"]), S.hv_coeff(params["dz"], params["dt"]),
              ctypes.byref(flag), int(steps), tune, stream_of(fields.device), kms)
    return kms[0], kms[1]


def tune_grid(params, fields, steps=2, register=True):
    """Pick (mode, x block, z segment) for this grid. Returns (tuning, table)
    where table maps each timed tuning to its (x_ms, z_ms)."""
    bufs = [fields.state.clone(), fields.state_tmp.clone(), fields.state_work.clone()]
    table = {}
    _time_grid(params, fields, S.DEFAULT_TUNING, 1, bufs)  # warm-up
    for xb, zs in itertools.zip_longest(X_BLOCKS, Z_SEGMENTS):
        t = (0, xb or X_BLOCKS[0], zs or Z_SEGMENTS[0])
        table[t] = _time_grid(params, fields, t, steps, bufs)
    best_x = min(X_BLOCKS, key=lambda xb: next(v[0] for t, v in table.items() if t[1] == xb))
    best_z = min(Z_SEGMENTS, key=lambda zs: next(v[1] for t, v in table.items() if t[2] == zs))
    fused = (0, best_x, best_z)
    table[fused] = _time_grid(params, fields, fused, steps, bufs)
    staged = (1, best_x, best_z)
    table[staged] = _time_grid(params, fields, staged, steps, bufs)
    best = fused if sum(table[fused]) <= sum(table[staged]) else staged
    if register:
        S.set_tuning(fields.nx, fields.nz, best)
    del bufs
    torch.cuda.synchronize(fields.device)
    return best, table


def tune_ensemble(ens, steps=2):
    """Pick the fused (x block, z segment) for an Ensemble (timed in place)
    and prepare it. Returns (tuning, table)."""
    table = {}

    def timed(t):
        ens.set_tuning(t)
        kms = (ctypes.c_double * 2)()
        ens.evolve(steps, kernel_ms=kms)
        return kms[0], kms[1]

    for xb, zs in itertools.zip_longest(X_BLOCKS, Z_SEGMENTS):
        t = (0, xb or X_BLOCKS[0], zs or Z_SEGMENTS[0])
        table[t] = timed(t)
    best_x = min(X_BLOCKS, key=lambda xb: next(v[0] for t, v in table.items() if t[1] == xb))
    best_z = min(Z_SEGMENTS, key=lambda zs: next(v[1] for t, v in table.items() if t[2] == zs))
    best = (0, best_x, best_z)
    ens.set_tuning(best)
    return best, table


def tune_tape(solver, steps=2):
    """Pick the fused (x block, z kernel) for an AdjointSolver's tape forward,
    where one launch covers every member (the best shape for a batch of
    small grids is not the one for the same grid alone). Each candidate is
    timed with HIP events around whole tape-forward calls, run on the
    solver's first tape as scratch, from its fields' state in every member;
    z first (x = 512), then x with the best z. Sets solver.tuning and
    returns (tuning, table of ms per call)."""
    f = solver.f
    steps = max(1, min(int(steps), solver.every))
    start = f.state.unsqueeze(0).expand(solver.bshape)
    table = {}

    def timed(t):
        solver.tuning = t
        solver._tape_run(start, 0, steps, store=False)  # warm-up
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        solver._tape_run(start, 0, steps, store=False)
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1)

    for zs in Z_SEGMENTS:
        table[(0, 512, zs)] = timed((0, 512, zs))
    best_z = min(Z_SEGMENTS, key=lambda zs: table[(0, 512, zs)])
    for xb in X_BLOCKS:
        if (0, xb, best_z) not in table:
            table[(0, xb, best_z)] = timed((0, xb, best_z))
    best = min(((0, xb, best_z) for xb in X_BLOCKS), key=lambda t: table[t])
    solver.tuning = best
    return best, table

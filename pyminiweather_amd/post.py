"""Diagnostics (mirrors pyminiweather.post)."""
from . import _lib
from ._lib import ptr, stream_of


def compute_stats_async(params, fields):
    """Launch the mass/energy reduction; returns the (2,) device tensor
    fields.stats_out = [mass, energy] without synchronising."""
    _lib.call("pmw_compute_stats", ptr(fields.state), ptr(fields.hy_dens_cell),
              ptr(fields.hy_dens_theta_cell), fields.nx, fields.nz,
              float(params["dx"]), float(params["dz"]), ptr(fields.stats_workspace),
              fields.stats_workspace.numel(), ptr(fields.stats_out),
              stream_of(fields.device))
    return fields.stats_out


def compute_stats(params, fields):
    """Total mass and total energy of the interior
    (pyminiweather/post/stats.py:8-35): sum(rho)*dx*dz and
    sum(rho*(u^2+w^2) + rho*cv*T)*dx*dz, with rho, theta, p, T built from the
    perturbations plus the hydrostatic background. Returns two floats."""
    out = compute_stats_async(params, fields).cpu()
    return float(out[0]), float(out[1])


def compute_solution_variables(params, fields):
    """rho', u, w, theta' of the interior as a (4, nz, nx) device tensor
    (pyminiweather/post/stats.py:38-69); output-path helper, not on the
    time-stepping path."""
    nx, nz, hs = params["nx"], params["nz"], params["hs"]
    s = fields.state[:, hs:nz + hs, hs:nx + hs]
    hyr = fields.hy_dens_cell[hs:nz + hs, None]
    hyt = fields.hy_dens_theta_cell[hs:nz + hs, None]
    out = s.new_zeros((fields.nvariables, nz, nx))
    out[0] = s[0]
    out[1] = s[1] / (hyr + s[0])
    out[2] = s[2] / (hyr + s[0])
    out[3] = (s[3] + hyt) / (hyr + s[0]) - hyt / hyr
    return out

"""State layout, constants and quadrature (mirrors pyminiweather.data).

``Fields`` keeps the reference's attribute names and shapes
(pyminiweather/data/fields.py:7-55) with every array a float64 torch tensor
on the GPU, plus a few device-side extras the fused kernels need.
"""
from dataclasses import dataclass
from enum import Enum
from typing import Optional

import numpy as np
import torch

from . import IDS  # noqa: F401  (re-exported for reference-style imports)


class Constants(float, Enum):
    """Physical constants (pyminiweather/data/constants.py:4-27)."""

    hv_beta = 0.05
    p0 = 1.0e5
    C0 = 27.5629410929725921310572974482
    gamma = 1.40027894002789400278940027894
    grav = 9.8
    cp = 1004.0
    cv = 717.0
    rd = 287
    pi = 3.14159265358979323846264338327
    theta0 = 300.0
    exner0 = 1.0


class GaussianQuadrature:
    """Immutable 3-point Gauss-Legendre rule on [0, 1]
    (pyminiweather/data/quadrature.py:4-71). Host-side constants."""

    def __init__(self):
        qp = np.array([0.112701665379258311482073460022, 0.5,
                       0.887298334620741688517926539980])
        qw = np.array([0.277777777777777777777777777779,
                       0.444444444444444444444444444444,
                       0.277777777777777777777777777779])
        for a in (qp, qw):
            a.setflags(write=False)
        self.__npoints = 3
        self.__qpoints = qp
        self.__qweights = qw
        self.__qweights_outer = np.outer(qw, qw)
        self.__qpoints_grid_x, self.__qpoints_grid_z = np.meshgrid(qp, qp)
        self.__qweights_grid_x, self.__qweights_grid_z = np.meshgrid(qw, qw)

    npoints = property(lambda self: self.__npoints)
    qpoints = property(lambda self: self.__qpoints)
    qweights = property(lambda self: self.__qweights)
    qweights_outer = property(lambda self: self.__qweights_outer)
    qpoints_grid_x = property(lambda self: self.__qpoints_grid_x)
    qpoints_grid_z = property(lambda self: self.__qpoints_grid_z)
    qweights_grid_x = property(lambda self: self.__qweights_grid_x)
    qweights_grid_z = property(lambda self: self.__qweights_grid_z)


Quadrature = GaussianQuadrature()


@dataclass
class Fields:
    """Persistent arrays of one grid (pyminiweather/data/fields.py:7-55).

    State buffers are (4, nz+4, nx+4) float64 on the GPU, variable-major with
    x fastest. Extras without a reference counterpart:

    * ``state_work``: third state buffer; the fused RK stages write stage 1
      there so that no kernel reads and writes the same buffer.
    * ``src_w``: (nz, nx) gravity-wave w-momentum source, ``ic_type ==
      "gravity"`` only (solve/source.py:41-50 precomputed once).
    * ``inj_rows``: (nz,) int32 inflow-row mask, ``ic_type == "injection"``.
    * ``stats_workspace`` / ``stats_out``: reduction scratch and result.
    """

    nx: int
    nz: int
    hs: int
    nvariables: int
    s: int
    state: torch.Tensor
    state_tmp: torch.Tensor
    flux: torch.Tensor
    tend: torch.Tensor
    vals_x: torch.Tensor
    d3_vals_x: torch.Tensor
    vals_z: torch.Tensor
    d3_vals_z: torch.Tensor
    hy_dens_cell: torch.Tensor
    hy_dens_theta_cell: torch.Tensor
    hy_dens_int: torch.Tensor
    hy_pressure_int: torch.Tensor
    hy_dens_theta_int: torch.Tensor
    fourth_order_kernel: torch.Tensor
    first_order_kernel: torch.Tensor
    shape: tuple
    state_work: Optional[torch.Tensor] = None
    src_w: Optional[torch.Tensor] = None
    inj_rows: Optional[torch.Tensor] = None
    stats_workspace: Optional[torch.Tensor] = None
    stats_out: Optional[torch.Tensor] = None

    @property
    def device(self):
        return self.state.device


def linspace(start, stop, num, device):
    """numpy.linspace(start, stop, num, endpoint=False), same rounding:
    i * ((stop - start) / num) + start."""
    step = (float(stop) - float(start)) / num
    return torch.arange(num, dtype=torch.float64, device=device) * step + float(start)


def initialize_fields(params, device=None, intermediates=True):
    """Allocate every array of ``Fields`` on the GPU, zero-filled
    (pyminiweather/data/fields.py:58-123). With ``intermediates=False`` the
    per-function scratch (flux, tend, vals_*, d3_vals_*) is left empty: the
    fused time step never touches it (used for ensemble members)."""
    from . import _lib

    nx, nz, hs, s = params["nx"], params["nz"], params["hs"], params["s"]
    assert hs * 2 - s == 0
    assert nx >= 2 and nz >= 2, "the 4-point stencil needs nx, nz >= 2"
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    nv = 4
    shape = (nv, nz + 2 * hs, nx + 2 * hs)

    def z(*shp):
        return torch.zeros(shp, dtype=torch.float64, device=device)

    def zi(*shp):
        return z(*shp) if intermediates else z(0)

    state = z(*shape)
    ic = params.get("ic_type", "thermal")
    src_w = z(nz, nx) if ic == "gravity" else None
    inj_rows = None
    if ic == "injection":
        # bcs.py:43-48: z = linspace(0, nz*dz, nz, endpoint=False) + 0.5
        zz = linspace(0.0, nz * params["dz"], nz, device) + 0.5
        zlen = params["zlen"]
        inj_rows = (torch.abs(zz - 3.0 * zlen / 4.0) <= zlen / 16.0).to(torch.int32)
    nws = int(_lib.load().pmw_stats_workspace_size(nx, nz))
    return Fields(
        nx=nx, nz=nz, hs=hs, nvariables=nv, s=s,
        state=state, state_tmp=state.clone(),
        flux=zi(nv, nz + 1, nx + 1), tend=zi(nv, nz, nx),
        vals_x=zi(nv, nz, nx + 1), d3_vals_x=zi(nv, nz, nx + 1),
        vals_z=zi(nv, nz + 1, nx), d3_vals_z=zi(nv, nz + 1, nx),
        hy_dens_cell=z(nz + 2 * hs), hy_dens_theta_cell=z(nz + 2 * hs),
        hy_dens_int=z(nz + 1), hy_pressure_int=z(nz + 1), hy_dens_theta_int=z(nz + 1),
        fourth_order_kernel=torch.tensor([-1.0 / 12, 7.0 / 12, 7.0 / 12, -1.0 / 12],
                                         dtype=torch.float64, device=device),
        first_order_kernel=torch.tensor([1.0, -3.0, 3.0, -1.0], dtype=torch.float64,
                                        device=device),
        shape=shape,
        state_work=z(*shape), src_w=src_w, inj_rows=inj_rows,
        stats_workspace=z(max(nws, 2)), stats_out=z(2),
    )

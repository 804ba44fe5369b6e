"""Initial and boundary conditions (mirrors pyminiweather.ics).

``set_bc_x`` / ``set_bc_z`` and ``init`` keep the reference's signatures
(pyminiweather/ics/bcs.py:8, 67; ics/initial.py:11) and run as HIP kernels.
"""
from enum import Enum

from . import _lib
from ._lib import ptr, stream_of

IC_KINDS = {"thermal": 0, "collision": 1, "gravity": 2, "density-current": 3,
            "injection": 4}
BC_PERIODIC, BC_INJECTION = 0, 1


class Directions(Enum):
    """Splitting direction (pyminiweather/ics/directions.py:4-6)."""

    X = 1
    Z = 2


def check_state(fields, t, name="state"):
    """A state buffer must be a contiguous float64 (4, nz+4, nx+4) tensor on
    the fields' device."""
    if t.shape != tuple(fields.shape) or t.dtype != fields.state.dtype \
            or not t.is_contiguous() or t.device != fields.state.device:
        raise ValueError(
            f"{name} must be a contiguous float64 tensor of shape {fields.shape} on "
            f"{fields.state.device}; got {tuple(t.shape)} {t.dtype} on {t.device}")


def bc_code(ic_type):
    if ic_type not in IC_KINDS:
        raise ValueError(f"unknown ic_type {ic_type!r}")
    return BC_INJECTION if ic_type == "injection" else BC_PERIODIC


def set_bc_x(params, fields, state_forcing, ic_type):
    """Periodic x halos on interior rows; injection inflow on the left for
    ic_type == "injection" (pyminiweather/ics/bcs.py:8-64)."""
    check_state(fields, state_forcing, "state_forcing")
    bc = bc_code(ic_type)
    _lib.call("pmw_set_bc_x", ptr(state_forcing), fields.nx, fields.nz, bc,
              ptr(fields.inj_rows) if bc == BC_INJECTION else None,
              ptr(fields.hy_dens_cell), ptr(fields.hy_dens_theta_cell),
              stream_of(fields.device))


def set_bc_z(params, fields, state_forcing, ic_type):
    """Rigid-wall z halos over all columns (pyminiweather/ics/bcs.py:67-148)."""
    check_state(fields, state_forcing, "state_forcing")
    bc_code(ic_type)
    _lib.call("pmw_set_bc_z", ptr(state_forcing), ptr(fields.hy_dens_cell),
              fields.nx, fields.nz, stream_of(fields.device))


def init(fields, params, Mesh, bubbles=None):
    """Fill the state and hydrostatic profiles for params["ic_type"]
    (pyminiweather/ics/initial.py:11-105), on the device.

    ``bubbles`` (optional, thermal kinds): sequence of up to 4
    (amplitude, x0, z0, xrad, zrad) cos^2 theta' bubbles replacing the
    reference's fixed ones; used by ensembles.
    """
    import ctypes

    ic_type = params["ic_type"]
    assert ic_type in IC_KINDS
    geo = (ctypes.c_double * 16)(*Mesh.linspace_geometry())
    if bubbles is not None:
        flat = [float(v) for b in bubbles for v in b]
        if len(flat) != 5 * len(bubbles) or not 0 < len(bubbles) <= 4:
            raise ValueError("bubbles must be 1..4 tuples of 5 numbers")
        bub = (ctypes.c_double * len(flat))(*flat)
        nb = len(bubbles)
    else:
        bub, nb = None, 0
    _lib.call("pmw_init", IC_KINDS[ic_type], ptr(fields.state), ptr(fields.hy_dens_cell),
              ptr(fields.hy_dens_theta_cell), ptr(fields.hy_dens_int),
              ptr(fields.hy_dens_theta_int), ptr(fields.hy_pressure_int),
              ptr(fields.src_w), fields.nx, fields.nz, geo, bub, nb,
              stream_of(fields.device))
    fields.state_tmp.copy_(fields.state)

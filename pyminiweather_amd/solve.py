"""Time stepping (mirrors pyminiweather.solve).

Reference-named entry points with the reference's signatures, each a thin
ctypes call into libpmw.so:

* ``interpolate_x/z``, ``compute_flux_x/z``, ``compute_tend_x/z``,
  ``add_source_terms``: one kernel each, filling the same ``Fields``
  intermediates as pyminiweather/solve/interpolate.py and source.py.
* ``discrete_step`` (solve/step.py:21-82): one fused kernel per call
  (halo + reconstruction + flux + tendency + source + RK update); when
  ``state_out`` overlaps ``state_forcing`` (the reference's aliased stage 2)
  it runs the fused tendency kernel into ``fields.tend`` and then the
  update, so the tendency is complete before the update, exactly like the
  reference (``fused=False`` runs the per-function kernels, which also fill
  the reference's intermediates).
* ``evolve`` (solve/step.py:85-143): one launch per direction, each running
  the three RK stages of that direction in LDS tiles (per-stage launches
  for the injection BC, or when chosen by the tuning). The first direction
  writes ``fields.state_work`` so no launch reads and writes one buffer;
  ``state`` and ``state_tmp`` end up exactly as the reference leaves them.

The fused paths do not materialise ``vals_*``, ``d3_vals_*``, ``flux`` or
``tend``; those are scratch in the reference too (every substep overwrites
them before reading).

miniWeather names are aliases: ``perform_timestep`` = ``evolve``,
``semi_discrete_step`` = ``discrete_step``, ``compute_tendencies_x/z``
(fused tendency into ``fields.tend``), ``set_halo_values_x/z`` =
``set_bc_x/z``.
"""
import ctypes

from . import _lib
from ._lib import ptr, stream_of
from .data import Constants
from .ics import BC_INJECTION, Directions, bc_code, check_state, set_bc_x, set_bc_z

# Direction toggle, a module global exactly like the reference's
# (solve/step.py:18): False -> Z then X; flipped after every evolve.
reverse_direction: bool = False


def hv_coeff(d, dt_full):
    """Hyperviscosity coefficient -hv_beta * d / (16 * dt) with the FULL
    timestep params["dt"] (solve/interpolate.py:99-101, 147-149)."""
    return -Constants.hv_beta.value * d / (16 * dt_full)


def _st(fields):
    return stream_of(fields.device)


def _inj(fields, ic_type):
    return ptr(fields.inj_rows) if bc_code(ic_type) == BC_INJECTION else None


def _src(params, fields):
    return ptr(fields.src_w) if params["ic_type"] == "gravity" else None


def interpolate_x(params, fields, state=None):
    """4th-order face values and 3rd differences along x into
    fields.vals_x / d3_vals_x (solve/interpolate.py:10-43)."""
    state = fields.state if state is None else state
    check_state(fields, state)
    _lib.call("pmw_interpolate_x", ptr(state), ptr(fields.vals_x), ptr(fields.d3_vals_x),
              fields.nx, fields.nz, _st(fields))


def interpolate_z(params, fields, state=None):
    """Same along z into fields.vals_z / d3_vals_z (interpolate.py:46-79)."""
    state = fields.state if state is None else state
    check_state(fields, state)
    _lib.call("pmw_interpolate_z", ptr(state), ptr(fields.vals_z), ptr(fields.d3_vals_z),
              fields.nx, fields.nz, _st(fields))


def compute_flux_x(params, fields):
    """x fluxes into fields.flux[:, :nz, :nx+1] (interpolate.py:82-129)."""
    _lib.call("pmw_compute_flux_x", ptr(fields.vals_x), ptr(fields.d3_vals_x),
              ptr(fields.hy_dens_cell), ptr(fields.hy_dens_theta_cell), ptr(fields.flux),
              fields.nx, fields.nz, hv_coeff(params["dx"], params["dt"]), _st(fields))


def compute_flux_z(params, fields):
    """z fluxes into fields.flux[:, :nz+1, :nx]; zeroes d3_vals_z[rho'] on the
    wall faces (interpolate.py:132-186)."""
    _lib.call("pmw_compute_flux_z", ptr(fields.vals_z), ptr(fields.d3_vals_z),
              ptr(fields.hy_dens_int), ptr(fields.hy_dens_theta_int),
              ptr(fields.hy_pressure_int), ptr(fields.flux), fields.nx, fields.nz,
              hv_coeff(params["dz"], params["dt"]), _st(fields))


def compute_tend_x(params, fields, state):
    """fields.tend = -(F[i+1]-F[i])/dx (interpolate.py:189-215)."""
    _lib.call("pmw_compute_tend_x", ptr(fields.flux), ptr(fields.tend), fields.nx,
              fields.nz, float(params["dx"]), _st(fields))


def compute_tend_z(params, fields, state):
    """fields.tend = -(G[k+1]-G[k])/dz - rho'(state)*g (interpolate.py:218-250)."""
    check_state(fields, state)
    _lib.call("pmw_compute_tend_z", ptr(fields.flux), ptr(state), ptr(fields.tend),
              fields.nx, fields.nz, float(params["dz"]), _st(fields))


def add_source_terms(params, mesh, fields):
    """Adds the gravity-wave forcing to the w tendency when
    ic_type == "gravity"; no-op otherwise (solve/source.py:53-75)."""
    if params["ic_type"] == "gravity":
        _lib.call("pmw_add_source_w", ptr(fields.tend), ptr(fields.src_w), fields.nx,
                  fields.nz, _st(fields))


def compute_tendencies_x(params, fields, state_forcing):
    """Fused set_bc_x + interpolate_x + compute_flux_x + compute_tend_x
    (+ source) into fields.tend."""
    check_state(fields, state_forcing, "state_forcing")
    if bc_code(params["ic_type"]) == BC_INJECTION:
        set_bc_x(params, fields, state_forcing, params["ic_type"])
        interpolate_x(params, fields, state_forcing)
        compute_flux_x(params, fields)
        compute_tend_x(params, fields, state_forcing)
        add_source_terms(params, None, fields)
        return
    _lib.call("pmw_tendencies_x", ptr(state_forcing), ptr(fields.hy_dens_cell),
              ptr(fields.hy_dens_theta_cell), _src(params, fields), ptr(fields.tend),
              fields.nx, fields.nz, float(params["dx"]),
              hv_coeff(params["dx"], params["dt"]), _st(fields))


def compute_tendencies_z(params, fields, state_forcing):
    """Fused set_bc_z + interpolate_z + compute_flux_z + compute_tend_z
    (+ source) into fields.tend."""
    check_state(fields, state_forcing, "state_forcing")
    _lib.call("pmw_tendencies_z", ptr(state_forcing), ptr(fields.hy_dens_cell),
              ptr(fields.hy_dens_int), ptr(fields.hy_dens_theta_int),
              ptr(fields.hy_pressure_int), _src(params, fields), ptr(fields.tend),
              fields.nx, fields.nz, float(params["dz"]),
              hv_coeff(params["dz"], params["dt"]), _st(fields))


def _overlap(a, b):
    a0, b0 = a.data_ptr(), b.data_ptr()
    a1 = a0 + a.numel() * a.element_size()
    b1 = b0 + b.numel() * b.element_size()
    return a0 < b1 and b0 < a1


def discrete_step(params, fields, mesh, state_init, state_forcing, state_out, dt,
                  direction, fused=True):
    """One RK stage in one direction (solve/step.py:21-82): halo of
    state_forcing, tendency from state_forcing, then
    state_out[interior] = state_init[interior] + dt * tend."""
    for name, t in (("state_init", state_init), ("state_forcing", state_forcing),
                    ("state_out", state_out)):
        check_state(fields, t, name)
    if direction not in (Directions.X, Directions.Z):
        raise ValueError(f"direction must be a Directions member, got {direction!r}")
    ic_type = params["ic_type"]
    if fused and not _overlap(state_out, state_forcing):
        dx_or_dz = params["dx"] if direction == Directions.X else params["dz"]
        _lib.call("pmw_rk_stage", 0 if direction == Directions.X else 1,
                  ptr(state_init), ptr(state_forcing), ptr(state_out),
                  ptr(fields.hy_dens_cell), ptr(fields.hy_dens_theta_cell),
                  ptr(fields.hy_dens_int), ptr(fields.hy_dens_theta_int),
                  ptr(fields.hy_pressure_int), _src(params, fields), bc_code(ic_type),
                  _inj(fields, ic_type), fields.nx, fields.nz, float(dx_or_dz), float(dt),
                  hv_coeff(dx_or_dz, params["dt"]), _st(fields))
        return
    if fused:  # out aliases forcing: the whole tendency first, then the update
        if direction == Directions.X:
            compute_tendencies_x(params, fields, state_forcing)
        else:
            compute_tendencies_z(params, fields, state_forcing)
        _lib.call("pmw_rk_update", ptr(state_init), ptr(fields.tend), ptr(state_out),
                  fields.nx, fields.nz, float(dt), _st(fields))
        return
    if direction == Directions.X:
        set_bc_x(params, fields, state_forcing, ic_type)
        interpolate_x(params, fields, state_forcing)
        compute_flux_x(params, fields)
        compute_tend_x(params, fields, state_forcing)
    else:
        set_bc_z(params, fields, state_forcing, ic_type)
        interpolate_z(params, fields, state_forcing)
        compute_flux_z(params, fields)
        compute_tend_z(params, fields, state_forcing)
    add_source_terms(params, mesh, fields)
    _lib.call("pmw_rk_update", ptr(state_init), ptr(fields.tend), ptr(state_out),
              fields.nx, fields.nz, float(dt), _st(fields))


# Kernel choice for evolve: (mode, fused x block size, fused z segment rows).
# mode 0 = one fused launch per direction, 1 = one launch per RK stage. The
# per-grid entries are filled by pyminiweather_amd.autotune.
DEFAULT_TUNING = (0, 512, 8096)
_tuning_table = {}


def tuning_for(nx, nz):
    return _tuning_table.get((nx, nz), DEFAULT_TUNING)


def set_tuning(nx, nz, tuning):
    _tuning_table[(nx, nz)] = tuple(int(v) for v in tuning)


def _evolve_n(params, fields, dt, nsteps):
    global reverse_direction
    flag = ctypes.c_int(1 if reverse_direction else 0)
    ic_type = params["ic_type"]
    tune = (ctypes.c_int * 3)(*tuning_for(fields.nx, fields.nz))
    _lib.call("pmw_evolve_ex", ptr(fields.state), ptr(fields.state_tmp),
              ptr(fields.state_work), ptr(fields.hy_dens_cell),
              ptr(fields.hy_dens_theta_cell), ptr(fields.hy_dens_int),
              ptr(fields.hy_dens_theta_int), ptr(fields.hy_pressure_int),
              _src(params, fields), bc_code(ic_type), _inj(fields, ic_type),
              fields.nx, fields.nz, float(params["dx"]), float(params["dz"]), float(dt),
              hv_coeff(params["dx"], params["dt"]), hv_coeff(params["dz"], params["dt"]),
              ctypes.byref(flag), int(nsteps), tune, _st(fields), None)
    reverse_direction = bool(flag.value)


def evolve(params, fields, mesh, dt: float = 1e-4) -> None:
    """One timestep: 2 directions x 3 RK stages (dt/3, dt/2, dt), Z first
    when reverse_direction is False; flips the toggle (solve/step.py:85-143)."""
    _evolve_n(params, fields, dt, 1)


def run(params, fields, mesh, nsteps, dt=None):
    """``nsteps`` consecutive ``evolve`` calls in one library call (no Python
    round trip per step). No reference counterpart; same result as the loop
    at pyminiweather/__main__.py:218-237 without its I/O."""
    _evolve_n(params, fields, params["dt"] if dt is None else dt, nsteps)


perform_timestep = evolve
semi_discrete_step = discrete_step
set_halo_values_x = set_bc_x
set_halo_values_z = set_bc_z

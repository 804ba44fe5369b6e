"""NumPy restatement of PyMiniWeather's hot path (CPU oracle, test-only).

Every function names the reference file:line it restates. Floating-point
operation order follows the reference exactly, so that on the same inputs the
oracle reproduces the reference bit for bit on the same CPU:

* reconstruction: ``scipy.signal.convolve(mode="same")`` with method "direct"
  (chosen by SciPy at every grid size used here) computes, for face ``i`` and
  the padded row ``s``, ``(((s[i]*c0 + s[i+1]*c1) + s[i+2]*c2) + s[i+3]*c3)``
  with ``c`` the reversed kernel; the slicing path uses the same chain.
* quadrature sums: ``ndarray.sum(axis=-1)`` over 3 entries is ``(a+b)+c``.

Two reconstruction variants exist: ``"scipy"`` calls ``scipy.signal.convolve``
the way the reference does (solve/interpolate.py:33-43, 69-79) and is what the
CPU baseline times; ``"slice"`` is the same arithmetic by slicing, used for
long parity runs. Both are bit-identical (checked in the CPU tests).
"""
import numpy as np

__all__ = [
    "CONST", "QPOINTS", "QWEIGHTS", "HS", "NVARS", "DENS", "UMOM", "WMOM", "RHOT",
    "IC_TYPES", "make_params", "mesh_int_ext", "mesh_cell_centers",
    "mesh_vertical_cell_edges", "mesh_vertical_cell_centers_int_ext",
    "hydro_const_theta", "hydro_const_bvfreq", "sample_ellipse_cosine",
    "ic_point_values", "init_state", "OracleSolver",
    "halo_x", "halo_z", "reconstruct_x", "reconstruct_z", "flux_x", "flux_z",
    "tend_x", "tend_z", "gravity_source", "stats", "solution_variables",
]

# Physical constants (pyminiweather/data/constants.py:13-27). rd is declared
# as the integer 287 in a float Enum, so its value is 287.0.
CONST = {
    "hv_beta": 0.05,
    "p0": 1.0e5,
    "C0": 27.5629410929725921310572974482,
    "gamma": 1.40027894002789400278940027894,
    "grav": 9.8,
    "cp": 1004.0,
    "cv": 717.0,
    "rd": 287.0,
    "pi": 3.14159265358979323846264338327,
    "theta0": 300.0,
    "exner0": 1.0,
}
# 3-point Gauss-Legendre rule on [0, 1] (pyminiweather/data/quadrature.py:13-29)
QPOINTS = np.array([0.112701665379258311482073460022, 0.5,
                    0.887298334620741688517926539980])
QWEIGHTS = np.array([0.277777777777777777777777777779,
                     0.444444444444444444444444444444,
                     0.277777777777777777777777777779])
HS = 2          # halo width (pyminiweather/__main__.py:91-99)
NVARS = 4
DENS, UMOM, WMOM, RHOT = 0, 1, 2, 3   # pyminiweather/__init__.py:14-18
IC_TYPES = ("thermal", "collision", "gravity", "density-current", "injection")

# Convolution kernels as stored by the reference (data/fields.py:94-97) and the
# reversed copies that the direct correlation applies.
_K4 = np.array([-1.0 / 12, 7.0 / 12, 7.0 / 12, -1.0 / 12])
_K1 = np.array([1.0, -3.0, 3.0, -1.0])
_K4R = _K4[::-1].copy()
_K1R = _K1[::-1].copy()


def make_params(nx, nz, ic_type="thermal", xlen=2.0e4, zlen=1.0e4, cfl=1.0,
                max_speed=500.0, dt=None):
    """Parameter dict as built by pyminiweather/__main__.py:160-195."""
    p = {"nx": int(nx), "nz": int(nz), "xlen": float(xlen), "zlen": float(zlen),
         "hs": HS, "s": 4, "ic_type": ic_type, "cfl": float(cfl),
         "max_speed": float(max_speed)}
    p["dx"] = p["xlen"] / p["nx"]
    p["dz"] = p["zlen"] / p["nz"]
    p["dt"] = (float(np.minimum(p["dx"], p["dz"]) * p["cfl"] / p["max_speed"])
               if dt is None else float(dt))
    return p


# ---------------------------------------------------------------- mesh.py ----
def mesh_int_ext(p):
    """Cell left edges incl. halos, as a meshgrid (mesh.py:25-51)."""
    nx, nz, dx, dz = p["nx"], p["nz"], p["dx"], p["dz"]
    x = np.linspace(-HS * dx, (nx + HS) * dx, nx + 2 * HS, endpoint=False)
    z = np.linspace(-HS * dz, (nz + HS) * dz, nz + 2 * HS, endpoint=False)
    return np.meshgrid(x, z)


def mesh_cell_centers(p):
    """Interior cell centres (mesh.py:53-75)."""
    x = np.linspace(p["dx"] / 2.0, p["xlen"] + p["dx"] / 2.0, p["nx"], endpoint=False)
    z = np.linspace(p["dz"] / 2.0, p["zlen"] + p["dz"] / 2.0, p["nz"], endpoint=False)
    return np.meshgrid(x, z)


def mesh_vertical_cell_edges(p):
    """z of the nz+1 interior cell edges (mesh.py:77-92)."""
    return np.linspace(0.0, (p["nz"] + 1) * p["dz"], p["nz"] + 1, endpoint=False)


def mesh_vertical_cell_centers_int_ext(p):
    """z of the nz+4 cell centres incl. halos (mesh.py:94-112)."""
    nz, dz = p["nz"], p["dz"]
    return np.linspace((-HS + 0.5) * dz, (nz + HS + 0.5) * dz, nz + 2 * HS,
                       endpoint=False)


# ------------------------------------------------- initial_conditions.py ----
def hydro_const_theta(z):
    """Constant-theta hydrostatic background (ics/initial_conditions.py:26-52)."""
    c = CONST
    ht = c["theta0"]
    exner = c["exner0"] - c["grav"] * z / (c["cp"] * c["theta0"])
    pres = c["p0"] * (exner ** (c["cp"] / c["rd"]))
    hr = ((pres / c["C0"]) ** (1.0 / c["gamma"])) / ht
    return hr, ht


def hydro_const_bvfreq(z, bv_freq0):
    """Constant Brunt-Vaisala background (ics/initial_conditions.py:54-81)."""
    c = CONST
    ht = c["theta0"] * np.exp(bv_freq0 * bv_freq0 / c["grav"] * z)
    exner = c["exner0"] - c["grav"] * c["grav"] / (c["cp"] * bv_freq0 * bv_freq0) \
        * (ht - c["theta0"]) / (ht * c["theta0"])
    pres = c["p0"] * (exner ** (c["cp"] / c["rd"]))
    hr = ((pres / c["C0"]) ** (1.0 / c["gamma"])) / ht
    return hr, ht


def sample_ellipse_cosine(x, z, amplitude, x0, z0, xrad, zrad):
    """cos^2 bump inside an ellipse, zero outside (utils/utils.py:5-51)."""
    pi = CONST["pi"]
    dist = np.sqrt(((x - x0) / xrad) ** 2 + ((z - z0) / zrad) ** 2) * pi / 2.0
    out = np.zeros(z.shape, dtype=np.float64)
    inside = dist <= pi / 2.0
    np.putmask(out, inside, amplitude * (np.cos(dist) ** 2.0))
    return out


def ic_point_values(ic_type, x, z, xlen, bubble=None):
    """Point values (r, u, w, t, hr, ht) of one IC at coordinates x, z.

    Restates the CCQ factories (ics/initial_conditions.py:96-134 collision,
    166-199 thermal, 202-233 gravity, 236-268 density-current, 271-302
    injection). ``bubble`` optionally overrides the thermal bubble as
    (amplitude, x0, z0, xrad, zrad); the reference hard-codes
    (3, xlen/2, 2000, 2000, 2000).
    """
    zeros = np.zeros(x.shape)
    if ic_type == "gravity":
        hr, ht = hydro_const_bvfreq(z, 0.02)
        return zeros, 15.0 * np.ones(x.shape), zeros.copy(), zeros.copy(), hr, ht
    hr, ht = hydro_const_theta(z)
    t = np.zeros(x.shape)
    if ic_type == "thermal":
        a, x0, z0, xr, zr = bubble if bubble is not None else \
            (3.0, xlen / 2, 2000.0, 2000.0, 2000.0)
        t += sample_ellipse_cosine(x, z, a, x0, z0, xr, zr)
    elif ic_type == "collision":
        t = sample_ellipse_cosine(x, z, 20.0, xlen / 2, 2000.0, 2000.0, 2000.0)
        t += sample_ellipse_cosine(x, z, -20.0, xlen / 2, 8000.0, 2000.0, 2000.0)
    elif ic_type == "density-current":
        t += sample_ellipse_cosine(x, z, -20.0, xlen / 2, 5000.0, 4000.0, 2000.0)
    elif ic_type != "injection":
        raise ValueError(f"unknown ic_type {ic_type!r}")
    return zeros, zeros.copy(), zeros.copy(), t, hr, ht


def _hydro_for(ic_type, z):
    if ic_type == "gravity":
        return hydro_const_bvfreq(z, 0.02)
    return hydro_const_theta(z)


def init_state(p, bubble=None):
    """Allocate and initialise every persistent array.

    Restates data/fields.py:58-123 (shapes) and ics/initial.py:11-105 (cell
    averages by 3x3 Gauss-Legendre quadrature, then the 1-D hydrostatic
    arrays). Returns a dict of float64 arrays.
    """
    nx, nz, dx, dz = p["nx"], p["nz"], p["dx"], p["dz"]
    ic = p["ic_type"]
    if ic not in IC_TYPES:
        raise ValueError(f"unknown ic_type {ic!r}")
    qx, qz = np.meshgrid(QPOINTS, QPOINTS)   # qx[a,b] = q[b], qz[a,b] = q[a]
    w2 = np.outer(QWEIGHTS, QWEIGHTS)
    xe, ze = mesh_int_ext(p)
    x = xe[:, :, None, None] + qx * dx
    z = ze[:, :, None, None] + qz * dz
    r, u, w, t, hr, ht = ic_point_values(ic, x, z, p["xlen"], bubble)

    def cell_avg(v):
        return np.multiply(v, w2).sum(axis=-1).sum(axis=-1)

    shape = (NVARS, nz + 2 * HS, nx + 2 * HS)
    state = np.zeros(shape)
    state[DENS] = cell_avg(r)
    state[UMOM] = cell_avg((r + hr) * u)
    state[WMOM] = cell_avg((r + hr) * w)
    state[RHOT] = cell_avg((r + hr) * (t + ht) - hr * ht)

    zc = mesh_vertical_cell_centers_int_ext(p)
    hr_c, ht_c = _hydro_for(ic, zc)
    hy_dens_cell = hr_c * QWEIGHTS.sum()
    hy_dens_theta_cell = np.multiply((ht_c * hr_c)[:, None], QWEIGHTS).sum(axis=-1)
    ze1 = mesh_vertical_cell_edges(p)
    hr_e, ht_e = _hydro_for(ic, ze1)
    out = {
        "state": state,
        "state_tmp": state.copy(),
        "flux": np.zeros((NVARS, nz + 1, nx + 1)),
        "tend": np.zeros((NVARS, nz, nx)),
        "vals_x": np.zeros((NVARS, nz, nx + 1)),
        "d3_vals_x": np.zeros((NVARS, nz, nx + 1)),
        "vals_z": np.zeros((NVARS, nz + 1, nx)),
        "d3_vals_z": np.zeros((NVARS, nz + 1, nx)),
        "hy_dens_cell": hy_dens_cell,
        "hy_dens_theta_cell": hy_dens_theta_cell,
        "hy_dens_int": np.array(hr_e, dtype=np.float64),
        "hy_dens_theta_int": hr_e * ht_e,
        "hy_pressure_int": CONST["C0"] * ((hr_e * ht_e) ** CONST["gamma"]),
    }
    return out


# ---------------------------------------------------------------- bcs.py ----
def halo_x(p, f, s):
    """Periodic x halos on interior rows (ics/bcs.py:8-39), plus the injection
    inflow (bcs.py:41-64) when ic_type == "injection"."""
    nx, nz = p["nx"], p["nz"]
    rows = slice(HS, nz + HS)
    s[:, rows, 0] = s[:, rows, nx]
    s[:, rows, 1] = s[:, rows, nx + 1]
    if p["ic_type"] != "injection":
        s[:, rows, nx + HS] = s[:, rows, HS]
        s[:, rows, nx + HS + 1] = s[:, rows, HS + 1]
        return
    z = np.linspace(start=0, stop=nz * p["dz"], num=nz, endpoint=False) + 0.5
    idx = np.nonzero(np.fabs(z - 3.0 * p["zlen"] / 4.0) <= p["zlen"] / 16.0)[0] + HS
    hyr, hyt = f["hy_dens_cell"][idx], f["hy_dens_theta_cell"][idx]
    for col in (0, 1):
        s[UMOM, idx, col] = (s[DENS, idx, col] + hyr) * 50.0
        s[RHOT, idx, col] = (s[DENS, idx, col] + hyr) * 298.0 - hyt


def halo_z(p, f, s):
    """Wall z halos over all nx+4 columns (ics/bcs.py:67-148): rho*w = 0,
    rho*u scaled by the hydrostatic density ratio as (a/b)*c, rho' and
    (rho*theta)' copied from the first/last interior row."""
    nz = p["nz"]
    hy = f["hy_dens_cell"]
    lo, hi = HS, nz + HS - 1
    s[WMOM, [0, 1, nz + HS, nz + HS + 1], :] = 0.0
    for row in (0, 1):
        s[UMOM, row, :] = s[UMOM, lo, :] / hy[lo] * hy[row]
    for row in (nz + HS, nz + HS + 1):
        s[UMOM, row, :] = s[UMOM, hi, :] / hy[hi] * hy[row]
    for v in (DENS, RHOT):
        s[v, 0, :] = s[v, lo, :]
        s[v, 1, :] = s[v, lo, :]
        s[v, nz + HS, :] = s[v, hi, :]
        s[v, nz + HS + 1, :] = s[v, hi, :]


# -------------------------------------------------------- interpolate.py ----
def _chain(a, c):
    """((((a0*c0) + a1*c1) + a2*c2) + a3*c3) elementwise."""
    return ((a[0] * c[0] + a[1] * c[1]) + a[2] * c[2]) + a[3] * c[3]


def reconstruct_x(p, s, method="slice"):
    """Face values and 3rd differences along x on interior rows
    (solve/interpolate.py:10-43). Returns (vals, d3), each (4, nz, nx+1)."""
    nx, nz = p["nx"], p["nz"]
    rows = s[:, HS:nz + HS, :]
    if method == "scipy":
        from scipy.signal import convolve
        vals = convolve(rows, _K4[None, None, :], mode="same")[:, :, 2:-1]
        d3 = convolve(rows, _K1[None, None, :], mode="same")[:, :, 2:-1]
        return vals, d3
    taps = [rows[:, :, j:j + nx + 1] for j in range(4)]
    return _chain(taps, _K4R), _chain(taps, _K1R)


def reconstruct_z(p, s, method="slice"):
    """Face values and 3rd differences along z on interior columns
    (solve/interpolate.py:46-79). Returns (vals, d3), each (4, nz+1, nx)."""
    nx, nz = p["nx"], p["nz"]
    cols = s[:, :, HS:nx + HS]
    if method == "scipy":
        from scipy.signal import convolve
        vals = convolve(cols, _K4[None, :, None], mode="same")[:, 2:-1, :]
        d3 = convolve(cols, _K1[None, :, None], mode="same")[:, 2:-1, :]
        return vals, d3
    taps = [cols[:, j:j + nz + 1, :] for j in range(4)]
    return _chain(taps, _K4R), _chain(taps, _K1R)


def _hv_coeff(d, dt_full):
    # solve/interpolate.py:99-101 and 147-149: uses the full timestep
    return -CONST["hv_beta"] * d / (16 * dt_full)


def flux_x(p, f, vals, d3, dt_full=None):
    """x fluxes on the (nz, nx+1) faces (solve/interpolate.py:82-129).
    The pressure is not de-hydrostatized in x (interpolate.py:116)."""
    nz = p["nz"]
    hv = _hv_coeff(p["dx"], p["dt"] if dt_full is None else dt_full)
    rho = vals[DENS] + f["hy_dens_cell"][HS:nz + HS, None]
    u = vals[UMOM] / rho
    w = vals[WMOM] / rho
    t = (vals[RHOT] + f["hy_dens_theta_cell"][HS:nz + HS, None]) / rho
    pres = CONST["C0"] * np.power(rho * t, CONST["gamma"])
    return np.stack([
        rho * u - hv * d3[DENS],
        rho * u ** 2 + pres - hv * d3[UMOM],
        rho * u * w - hv * d3[WMOM],
        rho * u * t - hv * d3[RHOT],
    ])


def flux_z(p, f, vals, d3, dt_full=None):
    """z fluxes on the (nz+1, nx) faces (solve/interpolate.py:132-186).
    w = 0 and d3[rho'] = 0 on the two wall faces; ``d3`` is modified in place
    as the reference does (interpolate.py:168-173)."""
    nz = p["nz"]
    hv = _hv_coeff(p["dz"], p["dt"] if dt_full is None else dt_full)
    rho = vals[DENS] + f["hy_dens_int"][0:nz + 1, None]
    u = vals[UMOM] / rho
    w = vals[WMOM] / rho
    t = (vals[RHOT] + f["hy_dens_theta_int"][0:nz + 1, None]) / rho
    pres = CONST["C0"] * np.power(rho * t, CONST["gamma"]) \
        - f["hy_pressure_int"][0:nz + 1, None]
    w[0, :] = 0.0
    w[nz, :] = 0.0
    d3[DENS, 0, :] = 0.0
    d3[DENS, nz, :] = 0.0
    return np.stack([
        rho * w - hv * d3[DENS],
        rho * w * u - hv * d3[UMOM],
        rho * w ** 2 + pres - hv * d3[WMOM],
        rho * w * t - hv * d3[RHOT],
    ])


def tend_x(p, flux):
    """-(F[i+1]-F[i])/dx (solve/interpolate.py:189-215); flux is (4,nz,nx+1)."""
    return -(flux[:, :, 1:] - flux[:, :, :-1]) / p["dx"]


def tend_z(p, flux, s_forcing):
    """-(G[k+1]-G[k])/dz minus rho'*g from the forcing state
    (solve/interpolate.py:218-250); flux is (4,nz+1,nx)."""
    nx, nz = p["nx"], p["nz"]
    tend = -(flux[:, 1:, :] - flux[:, :-1, :]) / p["dz"]
    tend[WMOM] -= s_forcing[DENS, HS:nz + HS, HS:nx + HS] * CONST["grav"]
    return tend


def gravity_source(p, f, tend):
    """Gravity-wave forcing added to the w tendency every substep
    (solve/source.py:20-50); active only for ic_type == "gravity"
    (source.py:53-75)."""
    nz = p["nz"]
    x, z = mesh_cell_centers(p)
    wpert = sample_ellipse_cosine(x, z, 0.01, p["xlen"] / 8, 1000.0, 500.0, 500.0)
    tend[WMOM] += wpert * f["hy_dens_cell"][HS:nz + HS, None]


# ---------------------------------------------------------------- step.py ----
class OracleSolver:
    """Owns one state (dict from ``init_state``) and the direction toggle.

    ``stage`` restates ``discrete_step`` (solve/step.py:21-82) and
    ``timestep`` restates ``evolve`` (solve/step.py:85-143); the toggle is
    per instance and starts Z-first, like the reference's module global
    (step.py:18) in a fresh process.
    """

    def __init__(self, params, fields=None, method="slice", bubble=None):
        self.p = dict(params)
        self.f = fields if fields is not None else init_state(self.p, bubble)
        self.method = method
        self.reverse_direction = False

    def stage(self, s_init, s_forc, s_out, dt_stage, direction):
        p, f = self.p, self.f
        nx, nz = p["nx"], p["nz"]
        if direction == "x":
            halo_x(p, f, s_forc)
            vals, d3 = reconstruct_x(p, s_forc, self.method)
            f["vals_x"][...] = vals
            f["d3_vals_x"][...] = d3
            f["flux"][:, 0:nz, 0:nx + 1] = flux_x(p, f, f["vals_x"], f["d3_vals_x"])
            f["tend"][...] = tend_x(p, f["flux"][:, 0:nz, 0:nx + 1])
        elif direction == "z":
            halo_z(p, f, s_forc)
            vals, d3 = reconstruct_z(p, s_forc, self.method)
            f["vals_z"][...] = vals
            f["d3_vals_z"][...] = d3
            f["flux"][:, 0:nz + 1, 0:nx] = flux_z(p, f, f["vals_z"], f["d3_vals_z"])
            f["tend"][...] = tend_z(p, f["flux"][:, 0:nz + 1, 0:nx], s_forc)
        else:
            raise ValueError(f"direction must be 'x' or 'z', got {direction!r}")
        if p["ic_type"] == "gravity":
            gravity_source(p, f, f["tend"])
        interior = (slice(None), slice(HS, nz + HS), slice(HS, nx + HS))
        s_out[interior] = s_init[interior] + dt_stage * f["tend"]

    def timestep(self, dt=None):
        dt = self.p["dt"] if dt is None else dt
        s, t = self.f["state"], self.f["state_tmp"]
        order = ("x", "z") if self.reverse_direction else ("z", "x")
        for d in order:
            self.stage(s, s, t, dt / 3, d)
            self.stage(s, t, t, dt / 2, d)
            self.stage(s, t, s, dt / 1, d)
        self.reverse_direction = not self.reverse_direction

    def run(self, nsteps, dt=None):
        for _ in range(nsteps):
            self.timestep(dt)

    def stats(self):
        return stats(self.p, self.f, self.f["state"])


# --------------------------------------------------------------- stats.py ----
def stats(p, f, s):
    """Total mass and energy over the interior (post/stats.py:8-35).
    Kinetic energy is rho*(u^2+w^2) with no 1/2, as in the reference."""
    nx, nz, dx, dz = p["nx"], p["nz"], p["dx"], p["dz"]
    c = CONST
    sl = (slice(HS, nz + HS), slice(HS, nx + HS))
    rho = s[DENS][sl] + f["hy_dens_cell"][HS:nz + HS, None]
    u = s[UMOM][sl] / rho
    w = s[WMOM][sl] / rho
    th = (s[RHOT][sl] + f["hy_dens_theta_cell"][HS:nz + HS, None]) / rho
    pres = c["C0"] * np.power(rho * th, c["gamma"])
    temp = th / np.power(c["p0"] / pres, c["rd"] / c["cp"])
    ke = rho * (u * u + w * w)
    ie = rho * c["cv"] * temp
    return rho.sum() * dx * dz, (ke + ie).sum() * dx * dz


def solution_variables(p, f, s):
    """Derived output variables (post/stats.py:38-69): rho', u, w, theta'."""
    nx, nz = p["nx"], p["nz"]
    sl = (slice(HS, nz + HS), slice(HS, nx + HS))
    hyr = f["hy_dens_cell"][HS:nz + HS, None]
    hyt = f["hy_dens_theta_cell"][HS:nz + HS, None]
    out = np.zeros((NVARS, nz, nx))
    out[DENS] = s[DENS][sl]
    out[UMOM] = s[UMOM][sl] / (hyr + s[DENS][sl])
    out[WMOM] = s[WMOM][sl] / (hyr + s[DENS][sl])
    out[RHOT] = (s[RHOT][sl] + hyt) / (hyr + s[DENS][sl]) - hyt / hyr
    return out

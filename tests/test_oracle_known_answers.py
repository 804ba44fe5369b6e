"""Pins of the oracle that do not come from the reference's golden vectors:
the reference's own known-answer tests (tests/unit/*), closed forms, and
invariants of the scheme. Ends with planted mistakes: each plausible slip
in the oracle must be caught by one of these pins or by the golden vectors."""
import numpy as np
import pytest
from scipy.signal import convolve2d

import oracle as O
from conftest import load_golden


# ---- tests/unit/test_constants.py:7-45 ------------------------------------
def test_constants_known_values():
    c = O.CONST
    assert c["hv_beta"] == 0.05 and c["p0"] == 1.0e5
    assert c["C0"] == 27.5629410929725921310572974482
    assert c["gamma"] == 1.40027894002789400278940027894
    assert c["grav"] == 9.8 and c["cp"] == 1004.0 and c["cv"] == 717.0
    assert c["rd"] == 287 and c["theta0"] == 300.0 and c["exner0"] == 1.0
    assert c["pi"] == 3.14159265358979323846264338327


def test_quadrature_known_values_and_exactness():
    assert np.allclose(O.QPOINTS, [0.112701665379258311482073460022, 0.5,
                                   0.887298334620741688517926539980])
    assert np.allclose(O.QWEIGHTS, [0.277777777777777777777777777779,
                                    0.444444444444444444444444444444,
                                    0.277777777777777777777777777779])
    # 3-point Gauss-Legendre on [0, 1] integrates x^k exactly for k <= 5
    for k in range(6):
        assert abs(np.sum(O.QWEIGHTS * O.QPOINTS ** k) - 1.0 / (k + 1)) < 1e-15


# ---- tests/unit/test_mesh.py:18-82 (closed forms) --------------------------
def test_mesh_closed_forms():
    p = O.make_params(64, 32)
    xx, zz = O.mesh_int_ext(p)
    assert np.allclose(xx[0], (np.arange(68) - 2) * p["dx"])
    assert np.allclose(zz[:, 0], (np.arange(36) - 2) * p["dz"])
    xc, zc = O.mesh_cell_centers(p)
    assert np.allclose(xc[0], (np.arange(64) + 0.5) * p["dx"])
    assert np.allclose(zc[:, 0], (np.arange(32) + 0.5) * p["dz"])
    assert np.allclose(O.mesh_vertical_cell_edges(p), np.arange(33) * p["dz"])
    assert np.allclose(O.mesh_vertical_cell_centers_int_ext(p),
                       (np.arange(36) - 1.5) * p["dz"])


# ---- tests/unit/test_sample_cosine.py:9-73 ---------------------------------
def test_sample_ellipse_cosine_single_point():
    p = O.make_params(200, 100)
    x, z = O.mesh_int_ext(p)
    d = O.sample_ellipse_cosine(x, z, 1.0, p["xlen"] / 2, p["zlen"] / 2, 1e-4, 1e-4)
    assert np.count_nonzero(d == 0.0) == d.size - 1


def test_sample_ellipse_cosine_profile():
    p = O.make_params(200, 100)
    x, z = O.mesh_int_ext(p)
    x0, z0, r = p["xlen"] / 2, p["zlen"] / 2, 1e4
    d = O.sample_ellipse_cosine(x, z, 1.0, x0, z0, r, r)
    rr = np.hypot(x - x0, z - z0) / r
    expect = np.where(rr <= 1.0, np.cos(rr * np.pi / 2) ** 2, 0.0)
    assert np.allclose(d, expect, rtol=0, atol=1e-14)
    assert np.count_nonzero(d) == np.count_nonzero(rr <= 1.0)
    assert d.max() <= 1.0 and d.min() >= 0.0


# ---- tests/unit/test_initial_conditions.py:13-82 + hydrostatic balance -----
def test_thermal_ic_has_only_theta_perturbation():
    p = O.make_params(200, 100)
    f = O.init_state(p)
    s = f["state"]
    assert np.all(s[O.DENS] == 0) and np.all(s[O.UMOM] == 0) and np.all(s[O.WMOM] == 0)
    assert np.count_nonzero(s[O.RHOT]) > 0


def test_hydrostatic_profiles_small_mean_gradient():
    z = np.linspace(0, 100.0, 25000)
    hr, ht = O.hydro_const_theta(z)
    assert O.hydro_const_theta(1.0)[1] == O.hydro_const_theta(100.0)[1]
    assert np.isclose(np.gradient(hr).mean(), 0.0, atol=1e-4)
    hr2, ht2 = O.hydro_const_bvfreq(z, 0.02)
    assert np.isclose(np.gradient(hr2).mean(), 0.0, atol=1e-4)
    assert np.isclose(np.gradient(ht2).mean(), 0.0, atol=1e-4)


@pytest.mark.parametrize("which", ["theta", "bvfreq"])
def test_hydrostatic_balance(which):
    """dp/dz = -rho g with p = C0 (rho theta)^gamma, and theta(z) closed form."""
    c = O.CONST
    z = np.linspace(0.0, 1.0e4, 2001)
    h = 1e-2
    f = (lambda q: O.hydro_const_theta(q)) if which == "theta" else \
        (lambda q: O.hydro_const_bvfreq(q, 0.02))

    def pres(q):
        hr, ht = f(q)
        return c["C0"] * (hr * ht) ** c["gamma"]

    dpdz = (pres(z + h) - pres(z - h)) / (2 * h)
    hr, ht = f(z)
    assert np.max(np.abs(dpdz + hr * c["grav"]) / (hr * c["grav"])) < 1e-6
    if which == "bvfreq":
        assert np.allclose(ht, 300.0 * np.exp(0.02 ** 2 / 9.8 * z), rtol=1e-14)
    # surface pressure is p0 (exner0 = 1 at z = 0)
    assert abs(pres(np.array([0.0]))[0] - c["p0"]) < 1e-9 * c["p0"]


# ---- tests/unit/test_interpolate.py:11-63 + polynomial exactness -----------
@pytest.mark.parametrize("direction", ["x", "z"])
def test_reconstruction_matches_convolve2d_on_arange(direction):
    p = O.make_params(200, 100)
    shape = (4, p["nz"] + 4, p["nx"] + 4)
    s = np.arange(np.prod(shape), dtype=np.float64).reshape(shape)
    k4 = np.array([-1.0 / 12, 7.0 / 12, 7.0 / 12, -1.0 / 12])
    k1 = np.array([1.0, -3.0, 3.0, -1.0])
    if direction == "x":
        vals, d3 = O.reconstruct_x(p, s)
        for v in range(4):
            rows = s[v, 2:p["nz"] + 2, :]
            assert np.allclose(convolve2d(rows, k4[None, :], mode="same")[:, 2:-1], vals[v])
            assert np.allclose(convolve2d(rows, k1[None, :], mode="same")[:, 2:-1], d3[v])
    else:
        vals, d3 = O.reconstruct_z(p, s)
        for v in range(4):
            cols = s[v, :, 2:p["nx"] + 2]
            assert np.allclose(convolve2d(cols, k4[:, None], mode="same")[2:-1, :], vals[v])
            assert np.allclose(convolve2d(cols, k1[:, None], mode="same")[2:-1, :], d3[v])


def test_reconstruction_polynomial_exactness():
    """Face i sits between padded cells i+1 and i+2: the 4th-order value of a
    linear profile is its midpoint value; the 3rd difference of i^3 is 6
    (taps -1, 3, -3, 1 on i..i+3) and of a quadratic is 0."""
    p = O.make_params(10, 6)
    i = np.arange(p["nx"] + 4, dtype=np.float64)
    s = np.zeros((4, p["nz"] + 4, p["nx"] + 4))
    s[0] += 2.0 * i + 1.0
    s[1] += i ** 2
    s[2] += i ** 3
    vals, d3 = O.reconstruct_x(p, s)
    f = np.arange(p["nx"] + 1, dtype=np.float64)
    assert np.allclose(vals[0], 2.0 * (f + 1.5) + 1.0, rtol=0, atol=1e-13)
    assert np.allclose(d3[1], 0.0, atol=1e-12)
    assert np.allclose(d3[2], 6.0, rtol=0, atol=1e-10)


# ---- invariants of the full step -------------------------------------------
def test_hydrostatic_rest_stays_at_rest():
    p = O.make_params(40, 20, "thermal", cfl=1.5, max_speed=450.0)
    S = O.OracleSolver(p, bubble=(0.0, 1e4, 2e3, 2e3, 2e3))
    S.run(20)
    s = S.f["state"][:, 2:-2, 2:-2]
    hy = S.f["hy_dens_cell"][2:-2]
    assert np.max(np.abs(s[O.DENS])) < 1e-12 * hy.max()
    assert np.max(np.abs(s[O.UMOM])) < 1e-10 and np.max(np.abs(s[O.WMOM])) < 1e-10


@pytest.mark.parametrize("ic", ["thermal", "collision", "density-current"])
def test_mass_conserved(ic):
    p = O.make_params(40, 20, ic, cfl=1.5, max_speed=450.0)
    S = O.OracleSolver(p)
    m0, _ = S.stats()
    S.run(30)
    m1, _ = S.stats()
    assert abs(m1 - m0) / m0 < 1e-13


def test_mirror_symmetry():
    """Centred bubble: rho', rho*w, (rho*theta)' even in x about xlen/2,
    rho*u odd (to roundoff)."""
    p = O.make_params(40, 20, "collision", cfl=1.5, max_speed=450.0)
    S = O.OracleSolver(p)
    S.run(20)
    s = S.f["state"][:, 2:-2, 2:-2]
    for v, sign in ((O.DENS, 1), (O.UMOM, -1), (O.WMOM, 1), (O.RHOT, 1)):
        a = s[v]
        assert np.max(np.abs(a - sign * a[:, ::-1])) <= 1e-11 * max(np.max(np.abs(a)), 1e-300)


def test_periodic_shift_equivariance():
    """Shifting the interior by m columns commutes with a timestep, bitwise."""
    p = O.make_params(40, 20, "collision", cfl=1.5, max_speed=450.0)
    A = O.OracleSolver(p)
    A.run(3)
    B = O.OracleSolver(p, fields={k: (v.copy() if isinstance(v, np.ndarray) else v)
                                  for k, v in A.f.items()})
    B.reverse_direction = A.reverse_direction
    m = 7
    for key in ("state", "state_tmp"):
        B.f[key][:, :, 2:-2] = np.roll(A.f[key][:, :, 2:-2], m, axis=2)
    A.run(2)
    B.run(2)
    np.testing.assert_array_equal(np.roll(A.f["state"][:, 2:-2, 2:-2], m, axis=2),
                                  B.f["state"][:, 2:-2, 2:-2])


# ---- conditioning: why the gravity-wave IC gets a wider GPU tolerance ----------
def _run_with_pow_noise(monkeypatch, ic, nx, nz, steps, rate, seed=1):
    import oracle.numpy_ref as R
    rng = np.random.default_rng(seed)
    power = np.power

    def noisy_power(a, b, *args, **kw):
        r = power(a, b, *args, **kw)
        if isinstance(r, np.ndarray) and r.size > 100:
            flip = rng.random(r.shape) < rate
            up = rng.random(r.shape) < 0.5
            r = np.where(flip, np.nextafter(r, np.where(up, np.inf, -np.inf)), r)
        return r
    p = O.make_params(nx, nz, ic, cfl=1.5, max_speed=450.0)
    with monkeypatch.context() as m:
        m.setattr(R.np, "power", noisy_power)
        S = O.OracleSolver(p)
        S.run(steps)
    return S.f["state"]


def test_gravity_ic_pow_conditioning(monkeypatch):
    """Last-bit changes to 1% of pow results move the gravity-wave rho',
    rho*w, (rho*theta)' by > 1e-11 relative after 10 steps (and rho*u, the
    background-flow field, by < 1e-13), while the collision IC stays
    < 1e-12: the reference's own arithmetic cannot support a 1e-12 bar on
    those three gravity fields against any pow that is not glibc's."""
    g = load_golden("evolve_gravity_40x20")
    s = _run_with_pow_noise(monkeypatch, "gravity", 40, 20, 10, 0.01)
    rel = [np.linalg.norm(s[v] - g["state_10"][v]) / np.linalg.norm(g["state_10"][v])
           for v in range(4)]
    assert rel[0] > 1e-11 and rel[2] > 1e-11 and rel[3] > 1e-11 and rel[1] < 1e-13
    assert max(rel) < 1e-9
    g = load_golden("evolve_collision_40x20")
    s = _run_with_pow_noise(monkeypatch, "collision", 40, 20, 25, 0.01)
    rel = [np.linalg.norm(s[v] - g["state_25"][v]) / np.linalg.norm(g["state_25"][v])
           for v in range(4)]
    assert max(rel) < 1e-12


# ---- planted mistakes: each must fail a pin ----------------------------------
def _golden_pins_match():
    """True if the oracle reproduces both the seeded-random-state
    discrete_step golden (both directions) and 2 collision timesteps."""
    g = load_golden("discrete_step_thermal_12x7")
    p = O.make_params(12, 7, "thermal", dt=float(g["dt_full"]))
    S = O.OracleSolver(p)
    for d in ("x", "z"):
        a, b, c = g["s_init"].copy(), g["s_forc"].copy(), g["s_out"].copy()
        S.stage(a, b, c, float(g["dt_stage"]), d)
        if not np.array_equal(c, g[f"{d}_out"]):
            return False
    g = load_golden("evolve_collision_40x20")
    p = O.make_params(40, 20, "collision", cfl=1.5, max_speed=450.0)
    S = O.OracleSolver(p)
    S.run(2)
    return np.array_equal(S.f["state"], g["state_2"])


def test_golden_pin_passes_unmodified():
    assert _golden_pins_match()


@pytest.mark.parametrize("mistake", [
    "hv_uses_stage_dt", "x_first", "z_flux_u_w_swapped", "halo_x_off_by_one",
    "gravity_from_init_not_forcing", "d3_kernel_not_reversed", "tend_sign",
    "wall_keeps_d3_dens",
])
def test_planted_mistake_is_caught(monkeypatch, mistake):
    import oracle.numpy_ref as R
    if mistake == "hv_uses_stage_dt":
        orig = R.OracleSolver.stage

        def stage(self, a, b, c, dt_stage, d):
            saved = self.p["dt"]
            self.p["dt"] = dt_stage
            try:
                orig(self, a, b, c, dt_stage, d)
            finally:
                self.p["dt"] = saved
        monkeypatch.setattr(R.OracleSolver, "stage", stage)
    elif mistake == "x_first":
        orig_ts = R.OracleSolver.timestep

        def timestep(self, dt=None):
            self.reverse_direction = not self.reverse_direction
            orig_ts(self, dt)
            self.reverse_direction = not self.reverse_direction
        monkeypatch.setattr(R.OracleSolver, "timestep", timestep)
    elif mistake == "z_flux_u_w_swapped":
        orig_fz = R.flux_z

        def flux_z(p, f, vals, d3, dt_full=None):
            v2 = vals.copy()
            v2[[R.UMOM, R.WMOM]] = vals[[R.WMOM, R.UMOM]]
            return orig_fz(p, f, v2, d3, dt_full)
        monkeypatch.setattr(R, "flux_z", flux_z)
    elif mistake == "halo_x_off_by_one":
        def halo_x(p, f, st):
            nx, nz = p["nx"], p["nz"]
            rows = slice(2, nz + 2)
            st[:, rows, 0] = st[:, rows, nx - 1]
            st[:, rows, 1] = st[:, rows, nx]
            st[:, rows, nx + 2] = st[:, rows, 2]
            st[:, rows, nx + 3] = st[:, rows, 3]
        monkeypatch.setattr(R, "halo_x", halo_x)
    elif mistake == "gravity_from_init_not_forcing":
        orig_stage = R.OracleSolver.stage
        orig_tz = R.tend_z
        holder = {}

        def tend_z(p, flux, s_forcing):
            return orig_tz(p, flux, holder.get("init", s_forcing))

        def stage(self, a, b, c, dt_stage, d):
            holder["init"] = a
            try:
                orig_stage(self, a, b, c, dt_stage, d)
            finally:
                holder.clear()
        monkeypatch.setattr(R.OracleSolver, "stage", stage)
        monkeypatch.setattr(R, "tend_z", tend_z)
    elif mistake == "d3_kernel_not_reversed":
        monkeypatch.setattr(R, "_K1R", R._K1.copy())
    elif mistake == "tend_sign":
        monkeypatch.setattr(R, "tend_x", lambda p, flux: (flux[:, :, 1:] - flux[:, :, :-1]) / p["dx"])
    elif mistake == "wall_keeps_d3_dens":
        orig_fz = R.flux_z

        def flux_z(p, f, vals, d3, dt_full=None):
            keep = d3[R.DENS, [0, p["nz"]], :].copy()
            G = orig_fz(p, f, vals, d3, dt_full)
            hv = R._hv_coeff(p["dz"], p["dt"])
            G[R.DENS, [0, p["nz"]], :] -= hv * keep
            return G
        monkeypatch.setattr(R, "flux_z", flux_z)
    assert not _golden_pins_match()

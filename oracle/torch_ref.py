"""FP64 torch restatement of the time step, differentiable by autograd.

TEST INFRASTRUCTURE ONLY (part of the oracle): it is the autograd reference
against which the hand-written HIP adjoint is checked. It evolves only the
interior (4, nz, nx): halo values are rebuilt from it at every stage exactly
as set_bc_x / set_bc_z would (pyminiweather/ics/bcs.py:8-148), so the
interior trajectory equals the reference's. Formulas follow
solve/interpolate.py:82-250 and solve/step.py:21-143 operation by operation.
"""
import torch

from .numpy_ref import CONST, DENS, RHOT, UMOM, WMOM

_W4 = (-1.0 / 12, 7.0 / 12, 7.0 / 12, -1.0 / 12)   # taps s0..s3 (reversed kernel)
_D3 = (-1.0, 3.0, -3.0, 1.0)


def _recon(taps):
    v = ((taps[0] * _W4[0] + taps[1] * _W4[1]) + taps[2] * _W4[2]) + taps[3] * _W4[3]
    d = ((taps[0] * _D3[0] + taps[1] * _D3[1]) + taps[2] * _D3[2]) + taps[3] * _D3[3]
    return v, d


def _pad_x(u):
    """(4, nz, nx) -> (4, nz, nx+4) with periodic halo columns."""
    return torch.cat([u[..., -2:], u, u[..., :2]], dim=-1)


def _pad_z(u, hy_cell):
    """(4, nz, nx) -> (4, nz+4, nx) with the wall halo rows of set_bc_z."""
    nz = u.shape[1]
    lo, hi = u[:, 0], u[:, nz - 1]

    def rows(edge, pe, phs):
        out = []
        for ph in phs:
            out.append(torch.stack([
                edge[DENS],
                edge[UMOM] / hy_cell[pe] * hy_cell[ph],
                torch.zeros_like(edge[WMOM]),
                edge[RHOT]]))
        return torch.stack(out, dim=1)
    bot = rows(lo, 2, (0, 1))
    top = rows(hi, nz + 1, (nz + 2, nz + 3))
    return torch.cat([bot, u, top], dim=1)


def tend_x(u, f, dx, hv):
    """x tendency of the interior state u (interpolate.py:10-43, 82-129, 189-215)."""
    nx = u.shape[2]
    p = _pad_x(u)
    taps = [p[..., j:j + nx + 1] for j in range(4)]
    vals, d3 = _recon(taps)
    nz = u.shape[1]
    hyr = f["hy_dens_cell"][2:nz + 2, None]
    hyt = f["hy_dens_theta_cell"][2:nz + 2, None]
    rho = vals[DENS] + hyr
    uu = vals[UMOM] / rho
    ww = vals[WMOM] / rho
    tt = (vals[RHOT] + hyt) / rho
    pres = CONST["C0"] * torch.pow(rho * tt, CONST["gamma"])
    F = torch.stack([
        rho * uu - hv * d3[DENS],
        rho * uu ** 2 + pres - hv * d3[UMOM],
        rho * uu * ww - hv * d3[WMOM],
        rho * uu * tt - hv * d3[RHOT]])
    return -(F[..., 1:] - F[..., :-1]) / dx


def tend_z(u, f, dz, hv):
    """z tendency incl. gravity (interpolate.py:46-79, 132-186, 218-250)."""
    nz = u.shape[1]
    p = _pad_z(u, f["hy_dens_cell"])
    taps = [p[:, j:j + nz + 1] for j in range(4)]
    vals, d3 = _recon(taps)
    hyr = f["hy_dens_int"][:, None]
    hyt = f["hy_dens_theta_int"][:, None]
    hyp = f["hy_pressure_int"][:, None]
    rho = vals[DENS] + hyr
    uu = vals[UMOM] / rho
    ww = vals[WMOM] / rho
    tt = (vals[RHOT] + hyt) / rho
    pres = CONST["C0"] * torch.pow(rho * tt, CONST["gamma"]) - hyp
    wall = torch.zeros(nz + 1, 1, dtype=torch.bool, device=u.device)
    wall[0] = True
    wall[nz] = True
    ww = torch.where(wall, torch.zeros_like(ww), ww)
    d3_0 = torch.where(wall, torch.zeros_like(d3[DENS]), d3[DENS])
    G = torch.stack([
        rho * ww - hv * d3_0,
        rho * ww * uu - hv * d3[UMOM],
        rho * ww ** 2 + pres - hv * d3[WMOM],
        rho * ww * tt - hv * d3[RHOT]])
    t = -(G[:, 1:] - G[:, :-1]) / dz
    grav = torch.zeros_like(t)
    grav[WMOM] = u[DENS] * CONST["grav"]
    return t - grav


def evolve(u, p, f, nsteps, reverse_direction=False, src_w=None):
    """nsteps timesteps of the interior state u (4, nz, nx); returns the new
    interior. Direction order and stage dts as solve/step.py:85-143."""
    dt = p["dt"]
    hvx = -CONST["hv_beta"] * p["dx"] / (16 * dt)
    hvz = -CONST["hv_beta"] * p["dz"] / (16 * dt)

    def T(s, d):
        t = tend_x(s, f, p["dx"], hvx) if d == "x" else tend_z(s, f, p["dz"], hvz)
        if src_w is not None:
            add = torch.zeros_like(t)
            add[WMOM] = src_w
            t = t + add
        return t

    rev = reverse_direction
    for _ in range(nsteps):
        for d in (("x", "z") if rev else ("z", "x")):
            s1 = u + dt / 3 * T(u, d)
            s2 = u + dt / 2 * T(s1, d)
            u = u + dt / 1 * T(s2, d)
        rev = not rev
    return u

"""Gradients through the time step (adjoint form).

``AdjointSolver`` runs ``nsteps`` forward timesteps of one grid with the fused
HIP kernels, keeping a checkpoint of the state every ``checkpoint_every``
steps (default ceil(sqrt(nsteps))), and returns dL/d(initial state) for a
cotangent of the final state by sweeping the segments backwards: each
segment is recomputed forward from its checkpoint, then every step is
reversed with the hand-written adjoint kernels (``pmw_adjoint_step``).
Memory is O(sqrt(N)) states; recompute is one extra forward run.

``EvolveFunction`` wraps the same as a ``torch.autograd.Function`` (initial
interior state -> final interior state), so it composes with torch code.

The assimilation problem of the adjoint form: the control is a per-cell
theta' field mapped linearly into the initial (rho*theta)' by the per-row
hydrostatic density, (rho*theta)'_{k,i} = hy_dens_cell[k] * theta'_{k,i}
(the first-order map of ics/initial.py:73-77 with rho' = 0); the loss is
L = 1/2 * ||(rho*theta)'(T) - target||^2 over the interior.

No reference counterpart: the reference only steps forward.
"""
import ctypes
import math

import torch

from . import _lib
from . import solve as S
from ._lib import ptr, stream_of

RHOT = 3


class AdjointSolver:
    def __init__(self, params, fields, nsteps, checkpoint_every=None, tuning=None):
        if params["ic_type"] == "injection":
            raise ValueError("the adjoint supports the periodic-x ICs only")
        self.p = dict(params)
        self.f = fields
        self.nsteps = int(nsteps)
        self.every = int(checkpoint_every or max(1, math.ceil(math.sqrt(self.nsteps))))
        self.tuning = tuple(tuning) if tuning is not None else S.tuning_for(fields.nx,
                                                                              fields.nz)
        dev = fields.device
        shape = fields.shape
        self._buf = [torch.zeros(shape, dtype=torch.float64, device=dev) for _ in range(3)]
        n = int(_lib.load().pmw_adjoint_workspace_size(fields.nx, fields.nz))
        self._ws = torch.zeros(n, dtype=torch.float64, device=dev)
        self._hvx = S.hv_coeff(self.p["dx"], self.p["dt"])
        self._hvz = S.hv_coeff(self.p["dz"], self.p["dt"])

    # -- forward ---------------------------------------------------------
    def _steps(self, state, tmp, work, first_step, n):
        """Advance `state` by n steps starting at step index first_step (the
        direction order of step k is Z-first for even k, as the reference's
        toggle gives in a fresh run)."""
        f = self.f
        flag = ctypes.c_int(first_step % 2)
        tune = (ctypes.c_int * 3)(*self.tuning)
        _lib.call("pmw_evolve_ex", ptr(state), ptr(tmp), ptr(work), ptr(f.hy_dens_cell),
                  ptr(f.hy_dens_theta_cell), ptr(f.hy_dens_int), ptr(f.hy_dens_theta_int),
                  ptr(f.hy_pressure_int), ptr(f.src_w) if f.src_w is not None else None, 0,
                  None, f.nx, f.nz, float(self.p["dx"]), float(self.p["dz"]),
                  float(self.p["dt"]), self._hvx, self._hvz, ctypes.byref(flag), int(n), tune,
                  stream_of(f.device), None)

    def forward(self, state0):
        """Returns (final state, checkpoints). state0 is a state-shaped tensor
        (only its interior matters)."""
        s, t, w = self._buf
        s.copy_(state0)
        t.copy_(state0)
        ckpts = []
        done = 0
        while done < self.nsteps:
            ckpts.append(s.clone())
            n = min(self.every, self.nsteps - done)
            self._steps(s, t, w, done, n)
            done += n
        return s.clone(), ckpts

    # -- backward --------------------------------------------------------
    def backward(self, lam_final, ckpts):
        """dL/d(initial state) (interior; state layout) given dL/d(final
        state) (interior of lam_final)."""
        f = self.f
        lam = lam_final.clone()
        s, t, w = self._buf
        nseg = len(ckpts)
        for seg in reversed(range(nseg)):
            start = seg * self.every
            n = min(self.every, self.nsteps - start)
            states = [ckpts[seg].clone()]
            s.copy_(ckpts[seg])
            t.copy_(ckpts[seg])
            for k in range(1, n):
                self._steps(s, t, w, start + k - 1, 1)
                states.append(s.clone())
            for k in reversed(range(n)):
                step = start + k
                _lib.call("pmw_adjoint_step", ptr(states[k]), ptr(lam), ptr(self._ws),
                          self._ws.numel(), ptr(f.hy_dens_cell), ptr(f.hy_dens_theta_cell),
                          ptr(f.hy_dens_int), ptr(f.hy_dens_theta_int),
                          ptr(f.hy_pressure_int),
                          ptr(f.src_w) if f.src_w is not None else None, f.nx, f.nz,
                          float(self.p["dx"]), float(self.p["dz"]), float(self.p["dt"]),
                          self._hvx, self._hvz, step % 2, stream_of(f.device))
        return lam

    # -- the theta' assimilation problem ------------------------------------
    def row_weight(self):
        nz = self.f.nz
        return self.f.hy_dens_cell[2:nz + 2]

    def state_from_theta(self, theta):
        """Initial state with (rho*theta)' = hy_dens_cell[k] * theta' and the
        other perturbations zero; theta is (nz, nx)."""
        s0 = torch.zeros(self.f.shape, dtype=torch.float64, device=self.f.device)
        s0[RHOT, 2:-2, 2:-2] = self.row_weight()[:, None] * theta
        return s0

    def loss_and_grad(self, theta, target):
        """L(theta) = 1/2 ||(rho*theta)'(T) - target||^2 and dL/dtheta."""
        final, ckpts = self.forward(self.state_from_theta(theta))
        resid = final[RHOT, 2:-2, 2:-2] - target
        loss = 0.5 * float(torch.sum(resid * resid))
        lam = torch.zeros_like(final)
        lam[RHOT, 2:-2, 2:-2] = resid
        lam0 = self.backward(lam, ckpts)
        return loss, lam0[RHOT, 2:-2, 2:-2] * self.row_weight()[:, None]

    def loss(self, theta, target):
        final, _ = self.forward(self.state_from_theta(theta))
        resid = final[RHOT, 2:-2, 2:-2] - target
        return 0.5 * float(torch.sum(resid * resid))


class EvolveFunction(torch.autograd.Function):
    """Interior state (4, nz, nx) -> interior state after solver.nsteps steps,
    with the HIP adjoint as its backward."""

    @staticmethod
    def forward(ctx, interior, solver):
        s0 = torch.zeros(solver.f.shape, dtype=torch.float64, device=interior.device)
        s0[:, 2:-2, 2:-2] = interior
        final, ckpts = solver.forward(s0)
        ctx.solver = solver
        ctx.ckpts = ckpts
        return final[:, 2:-2, 2:-2].clone()

    @staticmethod
    def backward(ctx, grad):
        solver = ctx.solver
        lam = torch.zeros(solver.f.shape, dtype=torch.float64, device=grad.device)
        lam[:, 2:-2, 2:-2] = grad
        lam0 = solver.backward(lam, ctx.ckpts)
        return lam0[:, 2:-2, 2:-2].clone(), None

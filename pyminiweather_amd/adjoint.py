"""Gradients through the time step (adjoint form).

``AdjointSolver`` advances ``members`` identical grids (same size, profiles
and source; one member by default) for ``nsteps`` timesteps with the fused
HIP kernels, all members in one launch per direction, keeping a checkpoint
of the states every ``checkpoint_every`` steps (default ceil(sqrt(nsteps))).
The backward sweep takes the segments in reverse: each is recomputed from
its checkpoint into a tape that holds every stage state of the segment
(``pmw_tape_forward``), then swept backwards with the hand-written adjoint
kernels (``pmw_tape_adjoint``: one launch per direction, all three RK
stages, for all members). Two tapes alternate between segments, so the
recompute of segment s-1 runs on a second stream while the adjoint of
segment s runs on the caller's stream (events order each adjoint after its
recompute, and each recompute after the adjoint that last read its tape).
Memory is O(sqrt(N)) states plus two segments' tapes (6 states per step);
the recompute is one extra forward run. With ``keep_tapes`` (by default as
many as fit in three quarters of the device memory free when the solver is
made, after its two recompute tapes) the
forward sweep writes the full tapes of its last segments as it goes, and
the backward sweep uses them without recomputing: HBM capacity traded for
the recompute (the forward sweep pays the stage-state stores on those
segments).

``EvolveFunction`` wraps the same as a ``torch.autograd.Function`` (initial
interior state -> final interior state), so it composes with torch code.

The assimilation problem: the control is a per-cell theta' field mapped
linearly into the initial (rho*theta)' by the per-row hydrostatic density,
(rho*theta)'_{k,i} = hy_dens_cell[k] * theta'_{k,i} (the first-order map of
ics/initial.py:73-77 with rho' = 0); the loss of a member is
L = 1/2 * ||(rho*theta)'(T) - target||^2 over the interior.

No reference counterpart: the reference only steps forward.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from . import solve as S
from ._lib import ptr, stream_of
from .ics import Directions, check_state

RHOT = 3
# adjoint kernels: x block threads (256, 384), z tile (columns * 1000 +
# threads: 4256, 4384, 2256, 8512)
ADJ_DEFAULT_TUNING = (256, 8512)
ADJ_TUNINGS = ((256, 4256), (384, 4384), (256, 2256), (256, 8512))


# ---- per-function adjoints ---------------------------------------------------
# Transposes of the reference's halo, tendency and stage functions, for
# callers that chain their own reverse pass. Cotangent buffers have the state
# layout (4, nz+4, nx+4), float64 on the fields' device.

def _periodic_only(params):
    if params["ic_type"] == "injection":
        raise ValueError("adjoints support the periodic-x ICs only")


def set_bc_x_adjoint(params, fields, lam):
    """Transpose of set_bc_x (ics/bcs.py:8-39), in place on lam: halo-column
    cotangents are added to the columns they were copied from, then zeroed."""
    _periodic_only(params)
    check_state(fields, lam, "lam")
    _lib.call("pmw_adjoint_set_bc_x", ptr(lam), fields.nx, fields.nz, stream_of(fields.device))


def set_bc_z_adjoint(params, fields, lam):
    """Transpose of set_bc_z (ics/bcs.py:67-148), in place on lam."""
    check_state(fields, lam, "lam")
    _lib.call("pmw_adjoint_set_bc_z", ptr(lam), ptr(fields.hy_dens_cell), fields.nx,
              fields.nz, stream_of(fields.device))


def _stage_adjoint(params, fields, direction, state_forcing, lam, scale, out):
    _periodic_only(params)
    check_state(fields, state_forcing, "state_forcing")
    check_state(fields, lam, "lam")
    check_state(fields, out, "out")
    d = params["dx"] if direction == 0 else params["dz"]
    _lib.call("pmw_adjoint_stage", direction, ptr(state_forcing), ptr(lam), float(scale),
              None, None, None, ptr(out), ptr(fields.hy_dens_cell),
              ptr(fields.hy_dens_theta_cell), ptr(fields.hy_dens_int),
              ptr(fields.hy_dens_theta_int), fields.nx, fields.nz, float(d),
              S.hv_coeff(d, params["dt"]), stream_of(fields.device))


def _as_state(fields, lam_tend):
    lam = torch.zeros(fields.shape, dtype=torch.float64, device=fields.device)
    lam[:, 2:-2, 2:-2] = lam_tend
    return lam


def compute_tendencies_x_adjoint(params, fields, state_forcing, lam_tend):
    """d<lam_tend, T_x(state_forcing)> / d state_forcing, where T_x is
    compute_tendencies_x (halo + reconstruction + flux + tendency): returns
    the interior cotangent as a state-layout buffer (halos zero; the halo
    transpose is folded in). lam_tend is (4, nz, nx)."""
    out = torch.zeros(fields.shape, dtype=torch.float64, device=fields.device)
    _stage_adjoint(params, fields, 0, state_forcing, _as_state(fields, lam_tend), 1.0, out)
    return out


def compute_tendencies_z_adjoint(params, fields, state_forcing, lam_tend):
    """The z counterpart (gravity term included; the gravity-wave source is
    constant and has no cotangent)."""
    out = torch.zeros(fields.shape, dtype=torch.float64, device=fields.device)
    _stage_adjoint(params, fields, 1, state_forcing, _as_state(fields, lam_tend), 1.0, out)
    return out


def discrete_step_adjoint(params, fields, state_forcing, lam_out, dt, direction):
    """Transpose of discrete_step (solve/step.py:21-82) linearised at
    state_forcing: out = init + dt * T(forcing) gives (lam_init, lam_forcing)
    = (lam_out, dt * T'(forcing)^T lam_out) on the interiors (state layout,
    halos zero)."""
    d = 0 if direction == Directions.X else 1
    lam_init = torch.zeros(fields.shape, dtype=torch.float64, device=fields.device)
    lam_init[:, 2:-2, 2:-2] = lam_out[:, 2:-2, 2:-2]
    lam_forcing = torch.zeros_like(lam_init)
    _stage_adjoint(params, fields, d, state_forcing, lam_out, dt, lam_forcing)
    return lam_init, lam_forcing


class AdjointSolver:
    def __init__(self, params, fields, nsteps, checkpoint_every=None, tuning=None, members=1,
                 adjoint_tuning=None, keep_tapes="auto"):
        if params["ic_type"] == "injection":
            raise ValueError("the adjoint supports the periodic-x ICs only")
        self.p = dict(params)
        self.f = fields
        self.members = int(members)
        self.nsteps = int(nsteps)
        if self.nsteps < 1:
            raise ValueError("nsteps must be >= 1")
        self.every = int(checkpoint_every or max(1, math.ceil(math.sqrt(self.nsteps))))
        if tuning is not None:
            tun = tuple(tuning)
        elif self.members > 1:
            # one launch covers every member: the shapes tuned for large launches
            tun = S.DEFAULT_TUNING
        else:
            tun = S.tuning_for(fields.nx, fields.nz)
        if tun[0] != 0:
            tun = S.DEFAULT_TUNING
        self.tuning = tun
        lib = _lib.load()
        slot = int(lib.pmw_tape_slot_len(fields.nx, fields.nz, self.members))
        if slot <= 0:
            raise ValueError("bad grid or member count")
        self.bshape = (self.members,) + tuple(fields.shape)
        dev = fields.device
        n = min(self.every, self.nsteps)
        nseg = -(-self.nsteps // self.every)
        tshape = (6 * n + 1,) + self.bshape
        self._tapes = [torch.zeros(tshape, dtype=torch.float64, device=dev)
                       for _ in range(min(nseg, 2))]
        self._tape = self._tapes[0]
        # tapes kept from the forward sweep for the last segments
        tape_bytes = 8 * math.prod(tshape)
        if keep_tapes == "auto":
            free, _ = torch.cuda.mem_get_info(dev)
            keep = int((free * 3 // 4) // tape_bytes)
        else:
            keep = int(keep_tapes)
        keep = max(0, min(keep, nseg))
        self._kept = {seg: torch.zeros(tshape, dtype=torch.float64, device=dev)
                      for seg in range(nseg - keep, nseg)}
        self._kept_owner = None  # the checkpoint list of the run the kept tapes hold
        # recompute stream (default priority), used when there are >= 2 segments
        self._side = torch.cuda.Stream(device=dev) if nseg > 1 else None
        self._ws = torch.zeros((1,) + self.bshape, dtype=torch.float64, device=dev)
        self.adjoint_tuning = tuple(adjoint_tuning or ADJ_DEFAULT_TUNING)
        self._adj_tune = (ctypes.c_int * 2)(*self.adjoint_tuning)
        self._hvx = S.hv_coeff(self.p["dx"], self.p["dt"])
        self._hvz = S.hv_coeff(self.p["dz"], self.p["dt"])

    # -- helpers -----------------------------------------------------------
    def _batch(self, x):
        return x.reshape(self.bshape)

    def _tape_run(self, start, first_step, n, store, tape=None):
        """Tape slot 0 <- start; n steps from step index first_step (the
        direction order of step k is Z-first for even k, as the reference's
        toggle gives in a fresh run), on the current stream. Returns the
        end-state slot (a view)."""
        f = self.f
        tape = self._tape if tape is None else tape
        tape[0].copy_(start)
        tune = (ctypes.c_int * 3)(*self.tuning)
        _lib.call("pmw_tape_forward", ptr(tape), tape.numel(), self.members,
                  ptr(f.hy_dens_cell), ptr(f.hy_dens_theta_cell), ptr(f.hy_dens_int),
                  ptr(f.hy_dens_theta_int), ptr(f.hy_pressure_int),
                  ptr(f.src_w) if f.src_w is not None else None, f.nx, f.nz,
                  float(self.p["dx"]), float(self.p["dz"]), float(self.p["dt"]), self._hvx,
                  self._hvz, first_step % 2, int(n), int(store), tune, stream_of(f.device))
        return tape[6 * n]

    # -- forward / backward ------------------------------------------------
    def forward(self, state0):
        """Returns (final states, checkpoints); state0 has the state shape
        (one member) or (members,) + state shape; only interiors matter."""
        s = self._batch(state0)
        ckpts = []
        done = 0
        while done < self.nsteps:
            seg = len(ckpts)
            ckpts.append(s.clone())
            n = min(self.every, self.nsteps - done)
            if seg in self._kept:
                s = self._tape_run(s, done, n, store=True, tape=self._kept[seg])
            else:
                s = self._tape_run(s, done, n, store=False)
            done += n
        self._kept_owner = ckpts
        return s.clone().reshape(state0.shape), ckpts

    @property
    def kept_segments(self):
        """Number of segments whose tapes the forward sweep keeps."""
        return len(self._kept)

    def _tape_of(self, seg, kept):
        return self._kept[seg] if seg in kept else self._tapes[seg % 2]

    def _recompute(self, ckpts, seg):
        start = seg * self.every
        n = min(self.every, self.nsteps - start)
        self._tape_run(ckpts[seg], start, n, store=True, tape=self._tapes[seg % 2])

    def _adjoint(self, lam, seg, kept=()):
        f = self.f
        start = seg * self.every
        n = min(self.every, self.nsteps - start)
        tape = self._tape_of(seg, kept)
        _lib.call("pmw_tape_adjoint", ptr(tape), tape.numel(), self.members,
                  ptr(lam), ptr(self._ws), self._ws.numel(), ptr(f.hy_dens_cell),
                  ptr(f.hy_dens_theta_cell), ptr(f.hy_dens_int),
                  ptr(f.hy_dens_theta_int), f.nx, f.nz, float(self.p["dx"]),
                  float(self.p["dz"]), float(self.p["dt"]), self._hvx, self._hvz,
                  start % 2, int(n), self._adj_tune, stream_of(f.device))

    def backward(self, lam_final, ckpts):
        """dL/d(initial states) (interiors; state layout) given dL/d(final
        states) (interior of lam_final). Ordered on the caller's stream: on
        return, that stream follows every launch made here."""
        lam = self._batch(lam_final).clone()
        nseg = len(ckpts)
        # the kept tapes hold the latest forward run: used only for its own
        # checkpoints (another run's are recomputed in full)
        kept = self._kept if ckpts is self._kept_owner else {}
        nrec = nseg - len(kept)  # segments 0 .. nrec-1 are recomputed
        if nseg == 1 or nrec <= 1:
            for seg in reversed(range(nseg)):
                if seg not in kept:
                    self._recompute(ckpts, seg)
                self._adjoint(lam, seg, kept)
            return lam.reshape(lam_final.shape)
        main = torch.cuda.current_stream(self.f.device)
        side = self._side
        side.wait_stream(main)  # checkpoints and lam are ready
        ready = {}

        def recompute_on_side(seg):
            with torch.cuda.stream(side):
                self._recompute(ckpts, seg)
                ready[seg] = torch.cuda.Event()
                ready[seg].record(side)

        # the two latest recomputed segments first; they overlap the sweep
        # of the kept ones
        recompute_on_side(nrec - 1)
        recompute_on_side(nrec - 2)
        for seg in reversed(range(nseg)):
            if seg in kept:
                self._adjoint(lam, seg, kept)
                continue
            main.wait_event(ready.pop(seg))
            self._adjoint(lam, seg, kept)
            if seg >= 2:  # tape seg % 2 is free once this adjoint is done
                done = torch.cuda.Event()
                done.record(main)
                side.wait_event(done)
                recompute_on_side(seg - 2)
        return lam.reshape(lam_final.shape)

    # -- the theta' assimilation problem ------------------------------------
    def row_weight(self):
        nz = self.f.nz
        return self.f.hy_dens_cell[2:nz + 2]

    def state_from_theta(self, theta):
        """Initial states with (rho*theta)' = hy_dens_cell[k] * theta' and
        the other perturbations zero; theta is (nz, nx) or (members, nz, nx)."""
        th = theta.reshape(self.members, self.f.nz, self.f.nx)
        s0 = torch.zeros(self.bshape, dtype=torch.float64, device=self.f.device)
        s0[:, RHOT, 2:-2, 2:-2] = self.row_weight()[:, None] * th
        return s0

    def _losses(self, final, target):
        resid = final[:, RHOT, 2:-2, 2:-2] - target.reshape(self.members, self.f.nz, self.f.nx)
        return resid, 0.5 * torch.sum(resid * resid, dim=(1, 2))

    def loss_and_grad(self, theta, target):
        """L(theta) = 1/2 ||(rho*theta)'(T) - target||^2 and dL/dtheta. For a
        2-d theta: (float, (nz, nx) tensor); batched: ((members,) tensors)."""
        final, ckpts = self.forward(self.state_from_theta(theta))
        resid, losses = self._losses(self._batch(final), target)
        lam = torch.zeros(self.bshape, dtype=torch.float64, device=self.f.device)
        lam[:, RHOT, 2:-2, 2:-2] = resid
        lam0 = self.backward(lam, ckpts)
        grad = (lam0[:, RHOT, 2:-2, 2:-2] * self.row_weight()[:, None]).reshape(theta.shape)
        return (float(losses[0]), grad) if theta.dim() == 2 else (losses, grad)

    def loss(self, theta, target):
        final, _ = self.forward(self.state_from_theta(theta))
        _, losses = self._losses(self._batch(final), target)
        return float(losses[0]) if theta.dim() == 2 else losses

    def final_rhot(self, theta):
        final, _ = self.forward(self.state_from_theta(theta))
        return self._batch(final)[:, RHOT, 2:-2, 2:-2].reshape(
            theta.shape if theta.dim() == 3 else (self.f.nz, self.f.nx))

    def assimilate(self, theta0, target, iters=20):
        """Steepest descent on every member's theta' at once: per iteration
        one batched forward + checkpointed adjoint for the gradient g, then
        the step alpha = <g, g> / ||J g||^2 (exact for the linearised
        problem; J g from one more batched forward along g). theta0 and
        target are (members, nz, nx). Returns (theta, losses per iteration
        as a (iters + 1, members) array)."""
        theta = theta0.clone()
        history = []
        m = self.members
        for _ in range(iters):
            losses, g = self.loss_and_grad(theta, target)
            history.append(losses.cpu().numpy())
            gmax = torch.amax(torch.abs(g).reshape(m, -1), dim=1).clamp_min(1e-300)
            h = (1e-3 / gmax)[:, None, None]
            base = self.final_rhot(theta)
            jg = (self.final_rhot(theta + h * g) - base) / h
            num = torch.sum((g * g).reshape(m, -1), dim=1)
            den = torch.sum((jg * jg).reshape(m, -1), dim=1).clamp_min(1e-300)
            theta = theta - (num / den)[:, None, None] * g
        history.append(self.loss(theta, target).cpu().numpy())
        return theta, np.stack(history)


class EvolveFunction(torch.autograd.Function):
    """Interior state(s) (4, nz, nx) or (members, 4, nz, nx) -> the same after
    solver.nsteps steps, with the HIP adjoint as its backward."""

    @staticmethod
    def forward(ctx, interior, solver):
        s0 = torch.zeros(solver.bshape, dtype=torch.float64, device=interior.device)
        s0[:, :, 2:-2, 2:-2] = interior.reshape(solver.members, 4, solver.f.nz, solver.f.nx)
        final, ckpts = solver.forward(s0)
        ctx.solver = solver
        ctx.ckpts = ckpts
        ctx.in_shape = interior.shape
        return final[:, :, 2:-2, 2:-2].reshape(interior.shape).clone()

    @staticmethod
    def backward(ctx, grad):
        solver = ctx.solver
        lam = torch.zeros(solver.bshape, dtype=torch.float64, device=grad.device)
        lam[:, :, 2:-2, 2:-2] = grad.reshape(solver.members, 4, solver.f.nz, solver.f.nx)
        lam0 = solver.backward(lam, ctx.ckpts)
        return lam0[:, :, 2:-2, 2:-2].reshape(ctx.in_shape).clone(), None

"""Ensembles of uneven grids advanced together (batched form).

Each member is an ordinary ``Fields`` (its own grid size, timestep, IC);
``Ensemble.evolve`` advances all of them with one launch per direction and
timestep over the concatenated tile lists of all members
(``pmw_ensemble_evolve``), which leaves every member
bit-identical to ``solve.evolve`` on that member alone. ``capture`` records
a whole number of timestep pairs in a HIP graph (torch.cuda.CUDAGraph) so
that small members are not bound by launch latency.

No reference counterpart: the reference runs one grid per process
(pyminiweather/__main__.py:183-248); members use its IC and step exactly.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import PmwMember, stream_of
from .data import initialize_fields
from .ics import init
from .mesh import MeshData
from .post import compute_stats

GRID_CHOICES = ((200, 100), (400, 200), (800, 400))


def member_params(nx, nz, ic_type="thermal", cfl=1.5, max_speed=450.0, xlen=2.0e4,
                  zlen=1.0e4):
    """params dict of one member, dt from its own CFL (pyminiweather/
    __main__.py:190-195)."""
    p = {"nx": int(nx), "nz": int(nz), "xlen": float(xlen), "zlen": float(zlen), "hs": 2,
         "s": 4, "ic_type": ic_type, "cfl": float(cfl), "max_speed": float(max_speed)}
    p["dx"] = p["xlen"] / p["nx"]
    p["dz"] = p["zlen"] / p["nz"]
    p["dt"] = float(np.minimum(p["dx"], p["dz"]) * p["cfl"] / p["max_speed"])
    return p


def mixed_thermal_members(n=256, seed=0, grids=GRID_CHOICES):
    """The seeded mixed-size thermal ensemble: per member, in this order, a
    grid drawn uniformly from `grids`, then bubble amplitude ~ U(1,5) K,
    centre x ~ U(6,14) km, centre z ~ U(1.5,3) km, radius ~ U(1.5,2.5) km.
    Returns (params list, bubble list)."""
    rng = np.random.default_rng(seed)
    params, bubbles = [], []
    for _ in range(n):
        nx, nz = grids[int(rng.integers(len(grids)))]
        amp = float(rng.uniform(1.0, 5.0))
        x0 = float(rng.uniform(6.0e3, 14.0e3))
        z0 = float(rng.uniform(1.5e3, 3.0e3))
        r = float(rng.uniform(1.5e3, 2.5e3))
        params.append(member_params(nx, nz))
        bubbles.append((amp, x0, z0, r, r))
    return params, bubbles


class Ensemble:
    def __init__(self, params_list, bubbles=None, device=None, tuning=(0, 1256, 8128)):
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = device
        self.params = [dict(p) for p in params_list]
        self.fields, self.meshes = [], []
        for q, p in enumerate(self.params):
            if p.get("ic_type") == "injection":
                raise ValueError("ensembles support the periodic-x ICs only")
            f = initialize_fields(p, device, intermediates=False)
            mesh = MeshData(p, device)
            b = None if bubbles is None or bubbles[q] is None else [bubbles[q]]
            init(f, p, mesh, bubbles=b)
            self.fields.append(f)
            self.meshes.append(mesh)
        n = len(self.fields)
        self.members = (PmwMember * n)()
        for q, (p, f) in enumerate(zip(self.params, self.fields)):
            m = self.members[q]
            m.state, m.state_tmp, m.state_work = (f.state.data_ptr(), f.state_tmp.data_ptr(),
                                                  f.state_work.data_ptr())
            m.hy_dens_cell = f.hy_dens_cell.data_ptr()
            m.hy_dens_theta_cell = f.hy_dens_theta_cell.data_ptr()
            m.hy_dens_int = f.hy_dens_int.data_ptr()
            m.hy_dens_theta_int = f.hy_dens_theta_int.data_ptr()
            m.hy_pressure_int = f.hy_pressure_int.data_ptr()
            m.src_w = f.src_w.data_ptr() if f.src_w is not None else None
            m.nx, m.nz = f.nx, f.nz
            m.dx, m.dz = float(p["dx"]), float(p["dz"])
            m.dt = m.dt_full = float(p["dt"])
        self.reverse_direction = False
        self.set_tuning(tuning)

    def __len__(self):
        return len(self.fields)

    @property
    def cells(self):
        return sum(f.nx * f.nz for f in self.fields)

    def set_tuning(self, tuning):
        """Prepare a fresh workspace for `tuning`. Graphs captured earlier
        keep (and hold a reference to) the workspace they were captured
        with, whose descriptors match their launches."""
        self.tuning = tuple(int(v) for v in tuning)
        self._tune = (ctypes.c_int * 3)(*self.tuning)
        nbytes = int(_lib.load().pmw_ensemble_workspace_size(len(self)))
        ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        _lib.call("pmw_ensemble_prepare", self.members, len(self), self._tune,
                  ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_of(self.device))
        self.workspace = ws

    def evolve(self, nsteps=1, kernel_ms=None):
        """`nsteps` timesteps of every member (each with its own dt)."""
        flag = ctypes.c_int(1 if self.reverse_direction else 0)
        _lib.call("pmw_ensemble_evolve", self.members, len(self), self._tune,
                  ctypes.c_void_p(self.workspace.data_ptr()), ctypes.byref(flag),
                  int(nsteps), stream_of(self.device), kernel_ms)
        self.reverse_direction = bool(flag.value)

    def capture(self, step_pairs=1):
        """HIP graph of 2*step_pairs timesteps (an even count keeps the
        direction toggle where it was). Replaying it advances every member
        by that many steps."""
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        saved = self.reverse_direction
        with torch.cuda.graph(g, stream=s):
            self.evolve(2 * step_pairs)
        torch.cuda.current_stream(self.device).wait_stream(s)
        assert self.reverse_direction == saved
        g.pmw_workspace = self.workspace  # the descriptors its launches read
        return g

    def stats(self):
        return [compute_stats(p, f) for p, f in zip(self.params, self.fields)]

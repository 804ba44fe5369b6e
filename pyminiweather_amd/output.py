"""Solution output in the reference's on-disk format.

pyminiweather/__main__.py:220-233 appends, every `output_freq` steps, the
full state (4, nz+4, nx+4) reshaped to (-1, nx+4) to `app_filename` and the
solution variables (4, nz, nx) (post/stats.py:38-69) reshaped to (-1, nx) to
`<name>_svars.<ext>`, both with numpy.savetxt(..., delimiter=","), the first
write truncating the files. `SnapshotWriter` keeps that format but takes the
snapshot asynchronously: `capture` enqueues the device-to-host copies into
pinned buffers on the compute stream and returns at once, so the caller can
enqueue the next timesteps; `flush` waits for the copies and formats the
text while the GPU keeps stepping.
"""
import numpy as np
import torch

from .post import compute_solution_variables


def svars_filename(app_filename):
    """The reference's `fname + "_svars." + append` (the name must hold
    exactly one dot, as there)."""
    fname, append = app_filename.split(".")
    return fname + "_svars." + append


class SnapshotWriter:
    def __init__(self, params, fields, app_filename):
        self.params = params
        self.fields = fields
        self.app_filename = app_filename
        self.svars_filename = svars_filename(app_filename)
        self.touched = False
        nx, nz = params["nx"], params["nz"]
        self._state = torch.empty(fields.shape, dtype=torch.float64, pin_memory=True)
        self._svars = torch.empty((fields.nvariables, nz, nx), dtype=torch.float64,
                                  pin_memory=True)
        self._event = None

    def capture(self):
        """Enqueue copies of the current state and solution variables."""
        self.flush()
        self._state.copy_(self.fields.state, non_blocking=True)
        self._svars.copy_(compute_solution_variables(self.params, self.fields),
                          non_blocking=True)
        self._event = torch.cuda.Event()
        self._event.record()

    def max_rho_theta(self):
        """max(state[3]) of the captured snapshot (halos included, as the
        reference's fields.state[3].max())."""
        self._event.synchronize()
        return float(self._state[3].max())

    def flush(self):
        """Write the pending snapshot (if any) in the reference's format."""
        if self._event is None:
            return
        self._event.synchronize()
        mode = "a" if self.touched else "w"
        state = self._state.numpy()
        svars = self._svars.numpy()
        with open(self.app_filename, mode) as out:
            np.savetxt(out, state.reshape(-1, state.shape[-1]), delimiter=",")
        with open(self.svars_filename, mode) as out:
            np.savetxt(out, svars.reshape(-1, svars.shape[-1]), delimiter=",")
        self.touched = True
        self._event = None

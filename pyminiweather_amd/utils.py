"""Utilities of the reference's utils package (pyminiweather/utils).

``TimedCodeBlock`` (timing.py:15-36) logs "label: seconds" like the
reference's, but synchronises the GPU on entry and exit so the elapsed time
covers the device work launched inside the block. ``sample_ellipse_cosine``
(utils.py:5-51) is the stand-alone bubble profile; the initial conditions
evaluate the same profile inside their HIP kernel.
"""
import logging
import time as _time

import numpy as np
import torch

logger = logging.getLogger("pyminiweather.log")


def time():
    """Microseconds, like the reference's NumPy-backend clock (timing.py:9-12)."""
    return _time.time_ns() / 1e3


class TimedCodeBlock:
    def __init__(self, label: str = "Elapsed time"):
        self.label = label

    def __enter__(self):
        if torch.cuda.is_available():
            torch.cuda.synchronize()
        self.elapsed_time = time()
        return self.elapsed_time

    def __exit__(self, type, value, traceback):
        if torch.cuda.is_available():
            torch.cuda.synchronize()
        self.elapsed_time = (time() - self.elapsed_time) / 1e6
        logger.info(f"{self.label}: {self.elapsed_time} s")


def sample_ellipse_cosine(x, z, amplitude, x0, z0, xrad, zrad):
    """amplitude * cos(d)^2 inside the ellipse centred at (x0, z0) with radii
    (xrad, zrad), where d = pi/2 * (normalised distance), and 0 outside, at
    the points (x, z) (pyminiweather/utils/utils.py:5-51). x and z are torch
    tensors (any device) or NumPy arrays; the result has z's shape."""
    from .data import Constants
    pi = Constants.pi.value
    if isinstance(z, torch.Tensor):
        dist = torch.sqrt(((x - x0) / xrad) ** 2 + ((z - z0) / zrad) ** 2) * pi / 2.0
        return torch.where(dist <= pi / 2.0, amplitude * (torch.cos(dist) ** 2.0),
                           torch.zeros_like(dist))
    dist = np.sqrt(((x - x0) / xrad) ** 2 + ((z - z0) / zrad) ** 2) * pi / 2.0
    out = np.zeros(z.shape, dtype=np.float64)
    inside = dist <= pi / 2.0
    np.putmask(out, inside, amplitude * (np.cos(dist) ** 2.0))
    return out

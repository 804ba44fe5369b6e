"""Timing helper (mirrors pyminiweather/utils/timing.py:15-36).

``TimedCodeBlock`` logs "label: seconds" like the reference's, but
synchronises the GPU on entry and exit so the elapsed time covers the
device work launched inside the block.
"""
import logging
import time as _time

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

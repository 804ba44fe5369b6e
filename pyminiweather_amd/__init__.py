"""MI355X-native PyMiniWeather time-stepping path.

Drop-in for the reference's solver API (pyminiweather.data / .ics / .mesh /
.solve / .post) with state held in float64 torch tensors on the GPU and every
operation of the time step running as a hand-written HIP kernel for gfx950
in ``libpmw.so`` (C ABI: include/pmw.h). There is no CPU fallback: calling a
solver function without the built library raises.

    from pyminiweather_amd.data import initialize_fields
    from pyminiweather_amd.mesh import MeshData
    from pyminiweather_amd.ics import init
    from pyminiweather_amd.solve import evolve
    from pyminiweather_amd.post import compute_stats
"""
from enum import IntEnum


class IDS(IntEnum):
    """Variable index order of the state (pyminiweather/__init__.py:14-18)."""

    DENS = 0
    UMOM = 1
    WMOM = 2
    RHOT = 3


backend = "hip-gfx950"

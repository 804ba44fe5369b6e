"""CPU oracle for the PyMiniWeather time-stepping path.

TEST INFRASTRUCTURE ONLY. This package is a NumPy restatement of the
reference's algorithm, used as the checker for the HIP path. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product package ``pyminiweather_amd`` never does, and never
falls back to it.

Pinning (what checks each function, beyond restating the reference):

* constants, quadrature, mesh coordinates: the reference's own known-answer
  tests (tests/unit/test_constants.py:7-45, tests/unit/test_mesh.py:18-82),
  restated as fixtures-in-code in ``tests/test_oracle_known_answers.py``.
* ``sample_ellipse_cosine``: tests/unit/test_sample_cosine.py:9-73 (single
  nonzero for a tiny radius; closed-form cos^2 profile).
* hydrostatic profiles: tests/unit/test_initial_conditions.py:44-82 plus the
  hydrostatic-balance closed form dp/dz = -rho g.
* reconstruction (``reconstruct_x/z``): tests/unit/test_interpolate.py:11-63
  semantics (SciPy ``convolve2d`` on an ``arange`` state) and a bitwise
  comparison between the SciPy path and the slicing path.
* ``init_state``, ``stage`` (discrete_step), ``timestep`` (evolve),
  ``compute_stats``: golden vectors written by ``tests/golden/make_golden.py``
  from the reference itself (every IC type; seeded perturbed states with
  distinct and aliased buffers; 1/2/10/25/1500-step trajectories).
* invariants: hydrostatic rest, mass conservation, mirror symmetry, periodic
  shift equivariance.
"""
from .numpy_ref import *  # noqa: F401,F403

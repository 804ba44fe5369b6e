"""ctypes binding of libpmw.so (the C ABI declared in include/pmw.h).

The library is loaded lazily on first use, after ``torch`` has been imported
so that libpmw binds to the HIP runtime torch already loaded (both use the
SONAME libamdhip64.so.7). A missing library is an error, never a fallback.
"""
import ctypes
import os

import torch  # noqa: F401  (must be loaded before libpmw, see above)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpmw.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_LL = ctypes.c_longlong
_IP = ctypes.POINTER(ctypes.c_int)


class PmwMember(ctypes.Structure):
    """struct pmw_member of include/pmw.h."""

    _fields_ = [("state", _P), ("state_tmp", _P), ("state_work", _P),
                ("hy_dens_cell", _P), ("hy_dens_theta_cell", _P), ("hy_dens_int", _P),
                ("hy_dens_theta_int", _P), ("hy_pressure_int", _P), ("src_w", _P),
                ("nx", _I), ("nz", _I), ("dx", _D), ("dz", _D), ("dt", _D),
                ("dt_full", _D)]


_MP = ctypes.POINTER(PmwMember)

# name -> (restype, argtypes); must match include/pmw.h
SIGNATURES = {
    "pmw_abi_version": (_I, []),
    "pmw_error_string": (ctypes.c_char_p, [_I]),
    "pmw_set_bc_x": (_I, [_P, _I, _I, _I, _P, _P, _P, _P]),
    "pmw_set_bc_z": (_I, [_P, _P, _I, _I, _P]),
    "pmw_interpolate_x": (_I, [_P, _P, _P, _I, _I, _P]),
    "pmw_interpolate_z": (_I, [_P, _P, _P, _I, _I, _P]),
    "pmw_compute_flux_x": (_I, [_P, _P, _P, _P, _P, _I, _I, _D, _P]),
    "pmw_compute_flux_z": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _D, _P]),
    "pmw_compute_tend_x": (_I, [_P, _P, _I, _I, _D, _P]),
    "pmw_compute_tend_z": (_I, [_P, _P, _P, _I, _I, _D, _P]),
    "pmw_add_source_w": (_I, [_P, _P, _I, _I, _P]),
    "pmw_rk_update": (_I, [_P, _P, _P, _I, _I, _D, _P]),
    "pmw_tendencies_x": (_I, [_P, _P, _P, _P, _P, _I, _I, _D, _D, _P]),
    "pmw_tendencies_z": (_I, [_P, _P, _P, _P, _P, _P, _P, _I, _I, _D, _D, _P]),
    "pmw_rk_stage": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P,
                          _I, _I, _D, _D, _D, _P]),
    "pmw_evolve": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _I, _I,
                        _D, _D, _D, _D, _D, _IP, _I, _P]),
    "pmw_evolve_ex": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _I, _I,
                           _D, _D, _D, _D, _D, _IP, _I, _IP, _P,
                           ctypes.POINTER(ctypes.c_double)]),
    "pmw_ensemble_workspace_size": (_LL, [_I]),
    "pmw_ensemble_prepare": (_I, [_MP, _I, _IP, _P, _LL, _P]),
    "pmw_ensemble_evolve": (_I, [_MP, _I, _IP, _P, _IP, _I, _P,
                                 ctypes.POINTER(ctypes.c_double)]),
    "pmw_adjoint_set_bc_x": (_I, [_P, _I, _I, _P]),
    "pmw_adjoint_set_bc_z": (_I, [_P, _P, _I, _I, _P]),
    "pmw_adjoint_stage": (_I, [_I, _P, _P, _D, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _D, _D, _P]),
    "pmw_tape_slot_len": (_LL, [_I, _I, _I]),
    "pmw_tape_forward": (_I, [_P, _LL, _I, _P, _P, _P, _P, _P, _P, _I, _I, _D, _D, _D, _D, _D,
                              _I, _I, _I, _P, _P]),
    "pmw_tape_adjoint": (_I, [_P, _LL, _I, _P, _P, _LL, _P, _P, _P, _P, _I, _I, _D, _D, _D, _D,
                              _D, _I, _I, _IP, _P]),
    "pmw_stats_workspace_size": (_LL, [_I, _I]),
    "pmw_compute_stats": (_I, [_P, _P, _P, _I, _I, _D, _D, _P, _LL, _P, _P]),
    "pmw_compute_stats_noalloc": (_I, [_P, _P, _P, _I, _I, _D, _D, _P, _P]),
    "pmw_init": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _I, _I, _P, _P, _I, _P]),
    "pmw_hydro_pressure": (_I, [_P, _P, _I, _P]),
}

_lib = None


class PMWError(RuntimeError):
    """A libpmw entry point returned a nonzero code."""

    def __init__(self, name, code, text):
        super().__init__(f"{name} failed with code {code}: {text}")
        self.code = code


def load():
    """Load libpmw.so and declare every signature. Raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python -m pyminiweather_amd.build`"
                " (the HIP path has no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def call(name, *args):
    """Call an int-returning entry point and raise PMWError on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise PMWError(name, rc, lib.pmw_error_string(rc).decode())
    return rc


def ptr(t):
    """Device address of a tensor (None passes NULL)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_of(device=None):
    """hipStream_t handle of torch's current stream on `device`."""
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)

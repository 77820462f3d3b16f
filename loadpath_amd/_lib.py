"""ctypes binding of liblpmi.so (include/loadpath_mi.h).

The library is built in-tree (``python -m loadpath_amd.build``). There is no
fallback: without the built library, or without a GPU, importing this module
raises, so nothing can silently run a CPU path in its place.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liblpmi.so"

LP_OK = 0
LP_NOT_CONVERGED = 1
LP_SINGULAR = 2
LP_ERR_HIP = -1
LP_ERR_ARG = -2
LP_ERR_STATE = -3

PREC = {"fp64": 0, "fp32": 1, "mixed": 2}
SMOOTH = {"chebyshev": 0, "jacobi": 1}
FILTER = {"top3d": 0, "padded": 1}


class Options(C.Structure):
    _fields_ = [
        ("rtol", C.c_double),
        ("maxiter", C.c_int),
        ("coarse_cells", C.c_int),
        ("degree", C.c_int),
        ("smoother", C.c_int),
        ("omega", C.c_double),
        ("precision", C.c_int),
        ("max_rhs", C.c_int),
        ("lmax_iters", C.c_int),
        ("max_levels", C.c_int),
    ]


class SimpStats(C.Structure):
    _fields_ = [
        ("compliance", C.c_double),
        ("change", C.c_double),
        ("volume", C.c_double),
        ("cg_iters", C.c_int),
        ("cg_iters_total", C.c_int),
        ("oc_evals", C.c_int),
        ("seconds", C.c_double),
        ("beta", C.c_double),
        ("move", C.c_double),
        ("step_fit", C.c_double),
        ("lam", C.c_double),
        ("vol_held", C.c_int),
        ("converged", C.c_int),
    ]


class LoopOpts(C.Structure):
    """lp_loop_opts (include/loadpath_mi.h)."""
    _fields_ = [
        ("beta_start", C.c_double),
        ("beta_max", C.c_double),
        ("beta_every", C.c_int),
        ("eta", C.c_double),
        ("erode", C.c_double),
        ("move", C.c_double),
        ("adaptive_move", C.c_int),
        ("tol", C.c_double),
        ("stall_rtol", C.c_double),
        ("stall_window", C.c_int),
        ("emin", C.c_double),
        ("penal", C.c_double),
        ("schedule", C.c_void_p),
        ("schedule_len", C.c_int),
    ]


class LinearSolveError(RuntimeError):
    pass


_lib = None

_P = C.c_void_p
_I = C.c_int
_D = C.c_double

_SIGS = {
    "lp_hex8_ke": (_I, [_D, _I, _P]),
    "lp_grid_create": (_I, [C.POINTER(_P), _I, _I, _I, _P, _P, C.POINTER(Options)]),
    "lp_grid_destroy": (None, [_P]),
    "lp_set_moduli": (_I, [_P, _P, _I]),
    "lp_estimate_lmax": (_I, [_P, _I, _P]),
    "lp_grid_info": (_I, [_P, _P, _P, _P, _P]),
    "lp_apply": (_I, [_P, _I, _I, _P, _P, _I, _I]),
    "lp_precondition": (_I, [_P, _P, _P, _I, _I]),
    "lp_solve": (_I, [_P, _I, _P, _P, _I, _I, _P, _P]),
    "lp_strain_energy": (_I, [_P, _I, _P, _P, _I]),
    "lp_element_product": (_I, [_P, _I, _P, _P, _P, _I]),
    "lp_stress_pnorm": (_I, [_P, _I, _P, _P, _P, _P, _I, _D, _D, _D, _D, _D, _P, _P, _P, _P, _I]),
    "lp_hex8_cb": (_I, [_D, _I, _I, _P]),
    "lp_level_data": (_I, [_P, _I, _I, _P]),
    "lp_prolong": (_I, [_P, _P, _I, _I, _I, _I]),
    "lp_restrict": (_I, [_P, _P, _I, _I, _I, _I]),
    "lp_simp_create": (_I, [C.POINTER(_P), _P, _I, _P, _P, _D, _D, _I, _D, _D, _D, _I]),
    "lp_simp_destroy": (None, [_P]),
    "lp_simp_iterate": (_I, [_P, _I, _P]),
    "lp_simp_get": (_I, [_P, _I, _P]),
    "lp_simp_set_x": (_I, [_P, _P]),
    "lp_simp_set_optimizer": (_I, [_P, _I]),
    "lp_simp_set_loop": (_I, [_P, C.POINTER(LoopOpts), _P, _P]),
    "lp_nodal_stress": (_I, [_I, _I, _I, _D, _D, _I, _P, _P, _D, _P, _P, _P, _P, _P, _I]),
    "lp_label26": (_I, [_I, _I, _I, _P, _P, _P]),
    "lp_label26_noalloc": (_I, [_I, _I, _I, _P, _P, _P]),
    "lp_simp_set_ties": (_I, [_P, _P, _I, _I]),
    "lp_grid_filter": (_I, [_I, _I, _I, _D, _D, _P, _P, _I]),
    "lp_simp_moduli": (_I, [C.c_longlong, _P, _D, _D, _P, _P, _I]),
    "lp_mask_rhs": (_I, [_P, _I, _P, _P, _I]),
    "lp_cells_apply": (_I, [_I, _I, _I, _I, _P, C.c_longlong, _P, _P, _P, _P, _P, _I, _I]),
    "lp_coarse_residual": (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I]),
    "lp_add_correction": (_I, [_P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I]),
    "lp_dot": (_I, [_P, _P, C.c_longlong, _I, _P]),
    "lp_cast": (_I, [_P, _I, _P, _I, C.c_longlong]),
    "lp_element_energy": (_I, [_P, _P, _P, C.c_longlong, C.c_longlong, _P, _P]),
    "lp_simp_set_reduction": (_I, [_P, _P, _P]),
    "lp_simp_loop_info": (_I, [_P, _P, _P, _P]),
    "lp_set_solve_tol": (_I, [_P, _D, _I]),
    "lp_material_average": (_I, [_I, _I, _I, _D, _D, _P, _P, _P]),
    "lp_stiffness_create": (_I, [C.POINTER(_P), _I, _I, _I, _P, _P, _P, C.POINTER(C.c_longlong),
                                 C.POINTER(C.c_longlong)]),
    "lp_stiffness_pattern": (_I, [_P, _P, _P]),
    "lp_stiffness_assemble": (_I, [_P, _P, _P, _I]),
    "lp_stiffness_destroy": (_I, [_P]),
    "lp_stiffness_size": (_I, [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "lp_band_create": (_I, [C.POINTER(_P), _P, _P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "lp_band_factor": (_I, [_P, _P, _I, _I]),
    "lp_band_factor_csr": (_I, [_P, _P, _I]),
    "lp_band_solve": (_I, [_P, _I, _P, _P, _I, _I, C.POINTER(C.c_double)]),
    "lp_band_destroy": (_I, [_P]),
    "lp_stress_averaged": (_I, [_I, _I, _I, _D, _D, _I, _D, _I, _P, _P, _P, _D, _D, _D, _P, _P, _I, _P, _P,
                                _P, _P]),
    "lp_project": (_I, [C.c_longlong, _P, _D, _D, _P, _P, _I]),
    "lp_simp_displacements": (_I, [_P, _P]),
    "lp_filter_apply": (_I, [_I, _I, _I, _D, _I, _P, _P, _I]),
    "lp_last_error": (C.c_char_p, []),
    "lp_version": (_I, []),
    "lp_time_apply": (_I, [_P, _I, _I, _I, _I, _P]),
    "lp_tune": (_I, [_P, _I, _I]),
    "lp_autotune": (_I, [_P, _I, _I, _P, _P]),
    "lp_profile": (_I, [_P, _I, _I, _I]),
    "lp_profile_stats": (_I, [_P, _P, _P, _P]),
    "lp_mma_create": (_I, [C.POINTER(_P), C.c_longlong, _I, _P, _P, _D, _D, _D, _D, _D]),
    "lp_mma_destroy": (_I, [_P]),
    "lp_mma_tighten": (_I, [_P, _D]),
    "lp_mma_set_move": (_I, [_P, _D]),
    "lp_mma_update": (_I, [_P, _P, _D, _P, _P, _P, _P, _I]),
    "lp_mma_get": (_I, [_P, _I, _P]),
}

EXPORTS = tuple(_SIGS)


def load_library(path: os.PathLike | None = None):
    """Load liblpmi.so (no GPU needed to load; calls need one)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise ImportError(f"{p} is not built: run `python -m loadpath_amd.build` (hipcc, gfx950)")
    # torch first: its HIP runtime is then the one in the process and the
    # library binds to it. Loaded the other way round, two HIP runtimes share
    # the process and the library's finds no device once torch has
    # initialised its own.
    import torch  # noqa: F401

    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def lib():
    return load_library()


def require_gpu():
    """The HIP path needs a GPU; there is no CPU fallback."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("loadpath_amd needs an AMD GPU (HIP): torch.cuda.is_available() is False")


def last_error() -> str:
    return lib().lp_last_error().decode(errors="replace")


def check(rc: int, what: str = ""):
    if rc == LP_OK:
        return rc
    msg = last_error()
    if rc == LP_NOT_CONVERGED:
        raise LinearSolveError(msg)
    if rc == LP_SINGULAR:
        raise RuntimeError(f"{what}: coarsest factor is singular ({msg})")
    raise RuntimeError(f"{what} failed ({rc}): {msg}")


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)

"""ctypes binding of libfen.so (the C-ABI declared in include/fen.h).

The library is the only compute path: importing this module on a machine
where libfen.so has not been built raises immediately, and every call
checks its status code.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfen.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "fen.h")

FEN_OK = 0
ERRORS = {-1: "FEN_ERR_ARG", -2: "FEN_ERR_UNSUPPORTED", -3: "FEN_ERR_HIP", -4: "FEN_ERR_NONFINITE",
          -5: "FEN_ERR_STATE"}

P, PHI, U, V, W, DVO_X, DVO_Y, DVO_Z, S_X, S_Y, S_Z, VOF, RHO, MU = range(14)
BC_PERIODIC, BC_WALL, BC_INFLOW, BC_OUTFLOW = range(4)
LEFT, RIGHT, BOTTOM, TOP, FRONT, BACK = range(6)
FP64, FP32_REFINE = 0, 1
SOLVERS = {0: "pp", 1: "pn", 2: "nn", 3: "ppp", 4: "ppn", 5: "npn", 6: "nnn"}


class FenError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libfen.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    c_int, c_double, c_size_t, vp = ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "fen_create": (c_int, [ctypes.POINTER(vp), c_int, c_int, c_int, c_int, c_double, c_double, c_double, dp, ip,
                               c_int]),
        "fen_destroy": (c_int, [vp]),
        "fen_last_error": (ctypes.c_char_p, [vp]),
        "fen_set_params": (c_int, [vp, c_double, c_double, dp, c_double, c_int]),
        "fen_set_timestep": (c_int, [vp, c_double, dp]),
        "fen_set_dt_o": (c_int, [vp, c_double]),
        "fen_get_scalars": (c_int, [vp, dp, dp, dp, dp, dp]),
        "fen_field_count": (c_int, [vp, c_int, ctypes.POINTER(c_size_t)]),
        "fen_upload": (c_int, [vp, c_int, dp, c_size_t]),
        "fen_download": (c_int, [vp, c_int, dp, c_size_t]),
        "fen_set_bc_plane": (c_int, [vp, c_int, c_int, dp, c_size_t]),
        "fen_set_bc_type": (c_int, [vp, c_int, c_int, c_int]),
        "fen_get_bc_type": (c_int, [vp, c_int, c_int, ip]),
        "fen_update_ghost_nodes": (c_int, [vp, c_int]),
        "fen_ns_step": (c_int, [vp, c_int, dp, dp, dp]),
        "fen_ns_run": (c_int, [vp, c_int, c_int, dp]),
        "fen_poisson_solve": (c_int, [vp, dp, c_int]),
        "fen_poisson_solve_noalloc": (c_int, [vp, dp, c_int]),
        "fen_add_advection": (c_int, [vp, dp, dp, dp]),
        "fen_format_status": (c_int, [vp, c_int, c_double, c_double, ctypes.c_char_p, c_size_t]),
        "fen_write_grid_json": (c_int, [vp, ctypes.c_char_p]),
        "fen_solver_kind": (c_int, [vp, ip]),
        "fen_grid_delta": (c_int, [vp, dp]),
        "fen_synchronize": (c_int, [vp]),
        "fen_profile_enable": (c_int, [vp, c_int]),
        "fen_profile_sample": (c_int, [vp, c_int]),
        "fen_profile_reset": (c_int, [vp]),
        "fen_profile_count": (c_int, []),
        "fen_profile_name": (ctypes.c_char_p, [c_int]),
        "fen_profile_read": (c_int, [vp, c_int, dp, ctypes.POINTER(ctypes.c_longlong), dp]),
        "fen_profile_predecessor": (c_int, [vp, c_int, ip]),
        "fen_adjoint_record": (c_int, [vp, c_int, c_int, c_int, dp]),
        "fen_adjoint_backward": (c_int, [vp, dp, dp, dp, dp, dp, dp, dp]),
        "fen_set_multiphase": (c_int, [vp, c_double, c_double, c_double, c_double, c_double, c_double]),
        "fen_update_material_properties": (c_int, [vp]),
        "fen_advect_vof": (c_int, [vp, c_double]),
        "fen_set_vof_x_first": (c_int, [vp, c_int]),
        "fen_get_vof_x_first": (c_int, [vp, ip]),
        "fen_status_text": (c_int, [c_int, c_double, c_double, c_double, c_double, ctypes.c_char_p, c_size_t]),
        "fen_grid_json_text": (c_int, [c_int, c_int, c_int, dp, c_double, c_double, c_double, ctypes.c_char_p,
                                       c_size_t]),
        "fen_case_read": (c_int, [ctypes.c_char_p, vp, ctypes.c_char_p, c_size_t]),
        "fen_save_state": (c_int, [vp, ctypes.c_char_p]),
        "fen_load_state": (c_int, [vp, ctypes.c_char_p]),
    }
    for name, (res, args) in sig.items():
        # a library built from older sources lacks the newest entry points:
        # they fail when called (tests/test_abi.py checks every declared symbol)
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def declared_symbols(header=HEADER):
    """Names of the functions declared in include/fen.h."""
    import re
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(fen_\w+)\s*\(", text, re.M)))

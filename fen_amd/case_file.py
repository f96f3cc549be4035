"""Case files: the JSON run description read by libfen's fen_case_read
(format in include/fen.h), and a Solver set up from one.

``read(path)`` returns the parsed ``FenCase`` (the C struct, member for
member) or raises ``CaseError`` with the reader's message;
``solver_from_case(cs)`` does what the Fortran layer's setup_case does
through the Python mirror: create the context, set the module variables,
upload the initial fields and return (solver, dt).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib

MAXLIST = 32
PATHLEN = 256
BC_NAMES = {_lib.BC_PERIODIC: "Periodic", _lib.BC_WALL: "Wall", _lib.BC_INFLOW: "Inflow",
            _lib.BC_OUTFLOW: "Outflow"}


class FenCase(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("n", ctypes.c_int * 3), ("bc", ctypes.c_int * 6),
        ("constant_cfl", ctypes.c_int), ("precision", ctypes.c_int), ("restart", ctypes.c_int),
        ("steps", ctypes.c_int), ("status_every", ctypes.c_int),
        ("n_save_state", ctypes.c_int), ("save_state", ctypes.c_int * MAXLIST),
        ("n_save_fields", ctypes.c_int), ("save_fields", ctypes.c_int * MAXLIST),
        ("write_final", ctypes.c_int), ("multiphase", ctypes.c_int),
        ("origin", ctypes.c_double * 3), ("L", ctypes.c_double * 3),
        ("density", ctypes.c_double), ("viscosity", ctypes.c_double), ("g", ctypes.c_double * 3),
        ("cfl", ctypes.c_double), ("dt", ctypes.c_double), ("U", ctypes.c_double),
        ("rho0", ctypes.c_double), ("rho1", ctypes.c_double), ("mu0", ctypes.c_double), ("mu1", ctypes.c_double),
        ("sigma", ctypes.c_double), ("iface", ctypes.c_double * 3),
        ("name", ctypes.c_char * 64), ("data_dir", ctypes.c_char * PATHLEN),
        ("restart_dir", ctypes.c_char * PATHLEN), ("init", (ctypes.c_char * PATHLEN) * 5),
    ]

    def init_file(self, which):
        """Initial-field file of "u", "v", "w", "p" or "vof" ("" if none)."""
        return self.init["uvwp".index(which) if which != "vof" else 4].value.decode()


class CaseError(ValueError):
    pass


def read(path):
    lib = _lib.load()
    cs = FenCase()
    err = ctypes.create_string_buffer(512)
    rc = lib.fen_case_read(os.fsencode(path), ctypes.byref(cs), err, len(err))
    if rc != _lib.FEN_OK:
        raise CaseError(err.value.decode())
    return cs


def solver_from_case(cs, base_dir="."):
    """Solver + dt for a parsed case (restart files are not read here)."""
    from .solver import Solver
    dim = cs.dim
    s = Solver(dim, tuple(cs.n), tuple(cs.L), bc=list(cs.bc), x0=tuple(cs.origin), precision=cs.precision)
    s.set_params(density=cs.density, viscosity=cs.viscosity, g=tuple(cs.g), cfl=cs.cfl,
                 constant_cfl=bool(cs.constant_cfl))
    nx, ny, nz = s.n
    for which, fid in (("u", _lib.U), ("v", _lib.V), ("w", _lib.W), ("p", _lib.P)):
        path = cs.init_file(which)
        if not path:
            continue
        a = np.fromfile(os.path.join(base_dir, path), dtype=np.float64)
        if a.size != nx * ny * nz:
            raise CaseError(f"{path}: {a.size} values, the grid has {nx * ny * nz} cells")
        g = np.zeros(s.shape(fid), order="F")
        k = slice(1, 2) if dim == 2 else slice(1, -1)
        g[1:-1, 1:-1, k] = a.reshape((nx, ny, nz), order="F")
        s.upload(fid, g)
        s.update_ghost_nodes(fid)
    if cs.U > 0:
        dt = s.set_timestep(cs.U)
    else:
        dt = cs.dt
        s.set_dt_o(dt)
    return s, dt

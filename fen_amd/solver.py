"""Python mirror of FEN's solver interface over libfen's C-ABI.

Names and argument meanings follow the reference's Fortran modules so that
tests read like FEN's own programs:

* ``Solver(...)``                 grid%setup + init_solver (src/grid.f90:67, src/solver.f90:34)
* ``set_timestep(U)``             set_timestep (src/navier_stokes.f90:623)
* ``navier_stokes_solver(step, dt)`` / ``advance_solution``  (src/navier_stokes.f90:50)
* ``solve_poisson(f, gl)``        solve_poisson procedure pointer (src/poisson.f90:21-29)
* ``update_ghost_nodes(field)``   scalar%update_ghost_nodes (src/scalar.f90:223)
* ``add_advection(R)``            add_advection (src/navier_stokes.f90:261)
* ``status_line(step, time, dt)`` print_navier_stokes_solver_status (src/navier_stokes.f90:734)
* ``adjoint_record`` / ``adjoint_backward``  the discrete adjoint (no reference counterpart)

Arrays are numpy float64 in Fortran order with FEN's ghosted shapes.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from ._lib import FenError, P, PHI, U, V, W, DVO_X, DVO_Y, DVO_Z, S_X, S_Y, S_Z  # noqa: F401


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Solver:
    def __init__(self, dim, n, L, bc=None, x0=(0.0, 0.0, 0.0), precision=_lib.FP64):
        self.lib = _lib.load()
        nx, ny, nz = n
        if dim == 2:
            nz = 1
        self.dim = dim
        self.n = (nx, ny, nz)
        if bc is None:
            bc = [_lib.BC_PERIODIC] * 6
        h = ctypes.c_void_p()
        x0a = (ctypes.c_double * 3)(*x0)
        bca = (ctypes.c_int * 6)(*bc)
        rc = self.lib.fen_create(ctypes.byref(h), dim, nx, ny, nz, L[0], L[1], L[2], x0a, bca, precision)
        if rc != _lib.FEN_OK:
            raise FenError(rc, "fen_create failed")
        self.h = h

    # -- lifecycle -------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.fen_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != _lib.FEN_OK:
            msg = self.lib.fen_last_error(self.h)
            raise FenError(rc, msg.decode() if msg else "")
        return rc

    # -- grid / parameters ----------------------------------------------
    @property
    def delta(self):
        d = ctypes.c_double()
        self._check(self.lib.fen_grid_delta(self.h, ctypes.byref(d)))
        return d.value

    @property
    def solver(self):
        k = ctypes.c_int()
        self._check(self.lib.fen_solver_kind(self.h, ctypes.byref(k)))
        return _lib.SOLVERS[k.value]

    def set_params(self, density=1.0, viscosity=1.0, g=(0.0, 0.0, 0.0), cfl=1.0, constant_cfl=False):
        ga = (ctypes.c_double * 3)(*g)
        self._check(self.lib.fen_set_params(self.h, density, viscosity, ga, cfl, int(constant_cfl)))

    def set_timestep(self, U):
        dt = ctypes.c_double()
        self._check(self.lib.fen_set_timestep(self.h, U, ctypes.byref(dt)))
        return dt.value

    def set_dt_o(self, dt_o):
        self._check(self.lib.fen_set_dt_o(self.h, dt_o))

    def scalars(self):
        vals = [ctypes.c_double() for _ in range(5)]
        self._check(self.lib.fen_get_scalars(self.h, *[ctypes.byref(v) for v in vals]))
        return dict(zip(["maxdiv", "maxcfl", "dt_o", "dt_visc", "dt_conv"], [v.value for v in vals]))

    # -- fields -----------------------------------------------------------
    def shape(self, field):
        nx, ny, nz = self.n
        return (nx + 2, ny + 2, nz + 2) if (field <= W or field >= _lib.VOF) else (nx, ny, nz)

    def count(self, field):
        c = ctypes.c_size_t()
        self._check(self.lib.fen_field_count(self.h, field, ctypes.byref(c)))
        return c.value

    def upload(self, field, arr):
        a = np.asfortranarray(arr, dtype=np.float64)
        if a.shape != self.shape(field):
            raise ValueError(f"shape {a.shape} != {self.shape(field)}")
        self._check(self.lib.fen_upload(self.h, field, _dp(a), a.size))

    def download(self, field):
        a = np.zeros(self.shape(field), order="F")
        self._check(self.lib.fen_download(self.h, field, _dp(a), a.size))
        return a

    def set_bc_plane(self, field, face, plane):
        a = np.asfortranarray(plane, dtype=np.float64)
        self._check(self.lib.fen_set_bc_plane(self.h, field, face, _dp(a), a.size))

    def set_bc_type(self, field, face, t):
        self._check(self.lib.fen_set_bc_type(self.h, field, face, t))

    def bc_type(self, field, face):
        t = ctypes.c_int()
        self._check(self.lib.fen_get_bc_type(self.h, field, face, ctypes.byref(t)))
        return t.value

    def update_ghost_nodes(self, field):
        self._check(self.lib.fen_update_ghost_nodes(self.h, field))

    # -- the time step ------------------------------------------------------
    def navier_stokes_solver(self, step, dt, want_checks=True):
        d = ctypes.c_double(dt)
        if want_checks:
            md, mc = ctypes.c_double(), ctypes.c_double()
            self._check(self.lib.fen_ns_step(self.h, step, ctypes.byref(d), ctypes.byref(md), ctypes.byref(mc)))
            return d.value, md.value, mc.value
        self._check(self.lib.fen_ns_step(self.h, step, ctypes.byref(d), None, None))
        return d.value, None, None

    advance_solution = navier_stokes_solver

    def run(self, step0, nsteps, dt):
        d = ctypes.c_double(dt)
        self._check(self.lib.fen_ns_run(self.h, step0, nsteps, ctypes.byref(d)))
        return d.value

    def adjoint_record(self, step0, nsteps, every, dt):
        """Run nsteps steps (as run()) keeping checkpoints every `every` steps."""
        d = ctypes.c_double(dt)
        self._check(self.lib.fen_adjoint_record(self.h, step0, nsteps, every, ctypes.byref(d)))
        return d.value

    def adjoint_backward(self, seed, want_p=False):
        """seed: dL/d(final u, v, w) as interior arrays; returns dL/d(initial u, v, w)
        (and dL/d(initial p) when want_p)."""
        s = [np.asfortranarray(a, dtype=np.float64) for a in seed]
        for a in s:
            if a.shape != self.n:
                raise ValueError(f"seed shape {a.shape} != {self.n}")
        g = [np.zeros(self.n, order="F") for _ in range(4 if want_p else 3)]
        self._check(self.lib.fen_adjoint_backward(self.h, _dp(s[0]), _dp(s[1]), _dp(s[2]), _dp(g[0]), _dp(g[1]),
                                                  _dp(g[2]), _dp(g[3]) if want_p else None))
        return g

    # -- two-phase mode (MF build), 2-D ---------------------------------------
    def set_multiphase(self, rho0, rho1, mu0, mu1, sigma=0.0, beta=1.0):
        """allocate_vof_fields + allocate_multiphase_fields and the phase
        properties (src/solver.f90:81-94); then upload VOF and call
        update_material_properties."""
        self._check(self.lib.fen_set_multiphase(self.h, rho0, rho1, mu0, mu1, sigma, beta))

    def update_material_properties(self):
        self._check(self.lib.fen_update_material_properties(self.h))

    def advect_vof(self, dt):
        self._check(self.lib.fen_advect_vof(self.h, dt))

    @property
    def x_first(self):
        """volume_of_fluid_mod's x_first (src/volume_of_fluid.f90:30): the
        vof's x sweep runs first at the next advection; each advection flips it."""
        v = ctypes.c_int(0)
        self._check(self.lib.fen_get_vof_x_first(self.h, ctypes.byref(v)))
        return bool(v.value)

    @x_first.setter
    def x_first(self, flag):
        self._check(self.lib.fen_set_vof_x_first(self.h, 1 if flag else 0))

    def synchronize(self):
        self._check(self.lib.fen_synchronize(self.h))

    # -- operators ---------------------------------------------------------
    def solve_poisson(self, f, gl):
        a = np.asfortranarray(f, dtype=np.float64).copy(order="F")
        self._check(self.lib.fen_poisson_solve(self.h, _dp(a), gl))
        return a

    def solve_poisson_noalloc(self, f, gl):
        """fen_poisson_solve_noalloc: the same solve, allocating nothing."""
        a = np.asfortranarray(f, dtype=np.float64).copy(order="F")
        self._check(self.lib.fen_poisson_solve_noalloc(self.h, _dp(a), gl))
        return a

    def add_advection(self, R=None):
        if R is None:
            R = [np.zeros(self.n, order="F") for _ in range(3)]
        R = [np.asfortranarray(r, dtype=np.float64).copy(order="F") for r in R]
        self._check(self.lib.fen_add_advection(self.h, _dp(R[0]), _dp(R[1]), _dp(R[2])))
        return R

    def status_line(self, step, time, dt):
        buf = ctypes.create_string_buffer(256)
        rc = self.lib.fen_format_status(self.h, step, time, dt, buf, 256)
        if rc not in (_lib.FEN_OK, -4):
            self._check(rc)
        return buf.value.decode()

    def save_state(self, path):
        """fen_save_state: the restart file of the resident state."""
        self._check(self.lib.fen_save_state(self.h, os.fsencode(path)))

    def load_state(self, path):
        """fen_load_state: the resident state from a restart file."""
        self._check(self.lib.fen_load_state(self.h, os.fsencode(path)))

    def write_grid_json(self, path):
        self._check(self.lib.fen_write_grid_json(self.h, path.encode()))

    # -- profiling ---------------------------------------------------------
    def profile(self, kernels=None):
        """Time the named kernels (all if None, none if an empty list) with HIP events."""
        if kernels is None:
            mask = -1
        else:
            names = [self.lib.fen_profile_name(k).decode() for k in range(self.lib.fen_profile_count())]
            mask = 0
            for kname in kernels:
                mask |= 1 << names.index(kname)
        self._check(self.lib.fen_profile_enable(self.h, mask))

    def profile_sample(self, every):
        """Time only every `every`-th step, with an event pair around each selected launch."""
        self._check(self.lib.fen_profile_sample(self.h, every))

    def profile_predecessor(self, kernel):
        names = [self.lib.fen_profile_name(k).decode() for k in range(self.lib.fen_profile_count())]
        k = ctypes.c_int()
        self._check(self.lib.fen_profile_predecessor(self.h, names.index(kernel), ctypes.byref(k)))
        return names[k.value] if k.value >= 0 else None

    def profile_reset(self):
        self._check(self.lib.fen_profile_reset(self.h))

    def profile_read(self):
        out = {}
        for k in range(self.lib.fen_profile_count()):
            ms, n, b = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
            self._check(self.lib.fen_profile_read(self.h, k, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b)))
            if n.value:
                out[self.lib.fen_profile_name(k).decode()] = {"ms": ms.value, "launches": n.value, "bytes": b.value}
        return out

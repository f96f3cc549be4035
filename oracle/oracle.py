"""ctypes wrapper of the CPU oracle (oracle/fen_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's CPU-baseline leg, always as the checker, never as the thing
measured or shipped.  The product library (fen_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

P, PHI, U, V, W, DVO_X, DVO_Y, DVO_Z, S_X, S_Y, S_Z, VOF, RHO, MU = range(14)
BC_PERIODIC, BC_WALL, BC_INFLOW, BC_OUTFLOW = range(4)
SOLVERS = {0: "pp", 1: "pn", 2: "nn", 3: "ppp", 4: "ppn", 5: "npn", 6: "nnn"}

_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 3 + [dp, ip]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_solver_kind.argtypes = [ctypes.c_void_p]
        L.orc_delta.argtypes = [ctypes.c_void_p]
        L.orc_delta.restype = ctypes.c_double
        L.orc_set_params.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, dp,
                                     ctypes.c_double, ctypes.c_int]
        L.orc_set_timestep.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_set_timestep.restype = ctypes.c_double
        L.orc_set_dt_o.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_get_dt_o.argtypes = [ctypes.c_void_p]
        L.orc_get_dt_o.restype = ctypes.c_double
        L.orc_get_dt_limits.argtypes = [ctypes.c_void_p, dp, dp]
        L.orc_field.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_field.restype = dp
        L.orc_field_len.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_field_len.restype = ctypes.c_size_t
        L.orc_bc_plane.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_bc_plane.restype = dp
        L.orc_bc_plane_len.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_bc_plane_len.restype = ctypes.c_size_t
        L.orc_bc_type.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_set_bc_type.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_update_ghost_nodes.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_ns_step.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, dp]
        L.orc_add_advection.argtypes = [ctypes.c_void_p, dp, dp, dp]
        L.orc_divergence.argtypes = [ctypes.c_void_p, dp, dp, dp, dp]
        L.orc_gradient.argtypes = [ctypes.c_void_p, dp, dp, dp, dp]
        L.orc_poisson_solve.argtypes = [ctypes.c_void_p, dp, ctypes.c_int]
        L.orc_fft_c2c.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int]
        L.orc_fft_r2c.argtypes = [ctypes.c_int, dp, dp, dp]
        L.orc_fft_c2r.argtypes = [ctypes.c_int, dp, dp, dp]
        L.orc_thomas.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_double, ctypes.c_double, dp, dp]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        L.orc_adjoint_step.argtypes = [ctypes.c_void_p, ctypes.c_double] + [dp] * 7
        L.orc_adjoint_step.restype = ctypes.c_int
        L.orc_set_multiphase.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 6
        L.orc_set_multiphase.restype = ctypes.c_int
        L.orc_update_material_properties.argtypes = [ctypes.c_void_p]
        L.orc_mf_set_timestep.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_mf_set_timestep.restype = ctypes.c_double
        L.orc_advect_vof.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_vof_reconstruct.argtypes = [ctypes.c_void_p]
        L.orc_vof_aux.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_vof_aux.restype = dp
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"])
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """One FEN solver instance (grid + navier_stokes_mod + poisson_mod state)."""

    def __init__(self, dim, n, L, bc=None, x0=(0.0, 0.0, 0.0)):
        nx, ny, nz = n
        lx, ly, lz = L
        if bc is None:
            bc = [BC_PERIODIC] * 6
        self.dim = dim
        self.n = (nx, ny, 1 if dim == 2 else nz)
        x0a = (ctypes.c_double * 3)(*x0)
        bca = (ctypes.c_int * 6)(*bc)
        self.h = lib().orc_create(dim, nx, ny, nz, lx, ly, lz, x0a, bca)
        if not self.h:
            raise ValueError("unsupported Poisson boundary combination")

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def solver(self):
        return SOLVERS[lib().orc_solver_kind(self.h)]

    @property
    def delta(self):
        return lib().orc_delta(self.h)

    def set_params(self, density=1.0, viscosity=1.0, g=(0.0, 0.0, 0.0), cfl=1.0, constant_cfl=False):
        ga = (ctypes.c_double * 3)(*g)
        lib().orc_set_params(self.h, density, viscosity, ga, cfl, int(constant_cfl))

    def set_timestep(self, U):
        return lib().orc_set_timestep(self.h, U)

    def set_dt_o(self, dt_o):
        lib().orc_set_dt_o(self.h, dt_o)

    @property
    def dt_o(self):
        return lib().orc_get_dt_o(self.h)

    def dt_limits(self):
        """(dt_visc, dt_conv) as set_timestep / update_timestep left them."""
        a, b = ctypes.c_double(0.0), ctypes.c_double(0.0)
        lib().orc_get_dt_limits(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def ghosted_shape(self):
        nx, ny, nz = self.n
        return (nx + 2, ny + 2, nz + 2)

    def field(self, fid):
        """Fortran-ordered numpy view of the field storage (ghosted for p..w)."""
        n = lib().orc_field_len(self.h, fid)
        ptr = lib().orc_field(self.h, fid)
        flat = np.ctypeslib.as_array(ptr, shape=(n,))
        shape = self.ghosted_shape() if (fid <= W or fid >= VOF) else self.n
        return flat.reshape(shape, order="F")

    def bc_plane(self, fid, face):
        n = lib().orc_bc_plane_len(self.h, face)
        ptr = lib().orc_bc_plane(self.h, fid, face)
        gx, gy, gz = self.ghosted_shape()
        shape = [(gy, gz), (gy, gz), (gx, gz), (gx, gz), (gx, gy), (gx, gy)][face]
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(shape, order="F")

    def bc_type(self, fid, face):
        return lib().orc_bc_type(self.h, fid, face)

    def set_bc_type(self, fid, face, t):
        lib().orc_set_bc_type(self.h, fid, face, t)

    def update_ghost_nodes(self, fid):
        lib().orc_update_ghost_nodes(self.h, fid)

    def step(self, step, dt):
        d = ctypes.c_double(dt)
        md = ctypes.c_double(0.0)
        mc = ctypes.c_double(0.0)
        lib().orc_ns_step(self.h, step, ctypes.byref(d), ctypes.byref(md), ctypes.byref(mc))
        return d.value, md.value, mc.value

    def add_advection(self):
        r = [np.zeros(self.n, order="F") for _ in range(3)]
        lib().orc_add_advection(self.h, *[_p(a) for a in r])
        return r

    def divergence(self, u, v, w):
        out = np.zeros(self.n, order="F")
        lib().orc_divergence(self.h, _p(u), _p(v), _p(w), _p(out))
        return out

    def gradient(self, s):
        r = [np.zeros(self.n, order="F") for _ in range(3)]
        lib().orc_gradient(self.h, _p(s), *[_p(a) for a in r])
        return r

    def adjoint_step(self, dt, ub, vb, wb, pb, dvob):
        """Adjoint of one step, in place on Fortran-ordered interior arrays;
        the context's u, v, w must hold the state the step started from."""
        arrs = [ub, vb, wb, pb] + list(dvob)
        for a in arrs:
            assert a.flags["F_CONTIGUOUS"] and a.dtype == np.float64 and a.shape == tuple(self.n)
        rc = lib().orc_adjoint_step(self.h, dt, *[_p(a) for a in arrs])
        if rc != 0:
            raise ValueError("adjoint step needs a 3-D triply periodic context")

    # ---- two-phase (MF) solver, 2-D
    def set_multiphase(self, rho0, rho1, mu0, mu1, sigma=0.0, beta=1.0):
        if lib().orc_set_multiphase(self.h, rho0, rho1, mu0, mu1, sigma, beta) != 0:
            raise ValueError("the two-phase solver is 2-D with Periodic or Wall faces")

    def update_material_properties(self):
        lib().orc_update_material_properties(self.h)

    def mf_set_timestep(self, U):
        return lib().orc_mf_set_timestep(self.h, U)

    def advect_vof(self, dt):
        lib().orc_advect_vof(self.h, dt)

    def vof_reconstruct(self):
        lib().orc_vof_reconstruct(self.h)

    def vof_aux(self, which):
        """0 h, 1 d, 2 curv, 3 norm x, 4 norm y, 5 l x, 6 l y (ghosted views)."""
        n = int(np.prod(self.ghosted_shape()))
        ptr = lib().orc_vof_aux(self.h, which)
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(self.ghosted_shape(), order="F")

    def poisson_solve(self, f, gl):
        """solve_poisson in place on a Fortran-ordered array with ghost level gl."""
        assert f.flags["F_CONTIGUOUS"] and f.dtype == np.float64
        lib().orc_poisson_solve(self.h, _p(f), gl)
        return f


def fft_c2c(x, sign):
    x = np.asarray(x, dtype=np.complex128)
    re = np.ascontiguousarray(x.real)
    im = np.ascontiguousarray(x.imag)
    lib().orc_fft_c2c(len(x), _p(re), _p(im), sign)
    return re + 1j * im


def fft_r2c(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x)
    re = np.zeros(n // 2 + 1)
    im = np.zeros(n // 2 + 1)
    lib().orc_fft_r2c(n, _p(x), _p(re), _p(im))
    return re + 1j * im


def fft_c2r(X, n):
    re = np.ascontiguousarray(np.real(X), dtype=np.float64)
    im = np.ascontiguousarray(np.imag(X), dtype=np.float64)
    out = np.zeros(n)
    lib().orc_fft_c2r(n, _p(re), _p(im), _p(out))
    return out


def thomas(a, b, c, lam_x, lam_y, r):
    re = np.ascontiguousarray(np.real(r), dtype=np.float64)
    im = np.ascontiguousarray(np.imag(r), dtype=np.float64)
    a, b, c = (np.ascontiguousarray(t, dtype=np.float64) for t in (a, b, c))
    lib().orc_thomas(len(re), _p(a), _p(b), _p(c), lam_x, lam_y, _p(re), _p(im))
    return re + 1j * im


def num_threads():
    return lib().orc_num_threads()


def set_num_threads(n):
    lib().orc_set_num_threads(n)

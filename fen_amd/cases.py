"""Initial fields of the benchmark and parity cases, written as FEN user
programs write them (interior values, then update_ghost_nodes).

The perturbations come from a counter-based SplitMix64 generator so that the
oracle and the GPU receive bit-identical inputs on any machine:
value(component c, Fortran linear index q) = mix(seed + (c * 2**40 + q + 1) * GOLDEN).
"""
from __future__ import annotations

import math

import numpy as np

from ._lib import BC_INFLOW, BC_OUTFLOW, BC_PERIODIC, BC_WALL, P, U, V, W

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
SEED = 20261202


def splitmix64(counter):
    z = counter.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(component, n, seed=SEED):
    """n uniforms in [0,1) for one component stream (53-bit mantissa)."""
    q = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        ctr = np.uint64(seed) + (np.uint64(component) * np.uint64(1 << 40) + q + np.uint64(1)) * GOLDEN
    return (splitmix64(ctr) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def gaussian(component, n, seed=SEED):
    """Box-Muller normals from two uniform streams (2c, 2c+1)."""
    u1 = uniform01(2 * component + 100, n, seed)
    u2 = uniform01(2 * component + 101, n, seed)
    u1 = np.where(u1 <= 0.0, 1e-300, u1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


class Case:
    """Grid, physics and initial interior fields of one run."""

    def __init__(self, name, dim, n, L, bc, density, viscosity, g, dt, dt_o, steps, fields, cfl=1.0, planes=None):
        self.name, self.dim, self.n, self.L, self.bc = name, dim, n, L, bc
        self.density, self.viscosity, self.g = density, viscosity, g
        self.dt, self.dt_o, self.steps, self.fields, self.cfl = dt, dt_o, steps, fields, cfl
        self.planes = planes or {}   # {(field, face): boundary values on the ghosted face plane}

    def apply(self, solver, upload):
        """Configure a solver (oracle or GPU) and write the initial fields.

        `upload(field, interior_array)` writes interior values; ghosts are then
        filled with update_ghost_nodes as FEN's init_fields routines do.
        """
        solver.set_params(density=self.density, viscosity=self.viscosity, g=self.g, cfl=self.cfl)
        solver.set_dt_o(self.dt_o)
        for (fid, face), plane in self.planes.items():
            if hasattr(solver, "set_bc_plane"):
                solver.set_bc_plane(fid, face, plane)
            else:
                solver.bc_plane(fid, face)[...] = plane
        for fid, arr in self.fields.items():
            upload(fid, arr)
        for fid in self.fields:
            solver.update_ghost_nodes(fid)


def ghosted(interior):
    """A ghosted Fortran-ordered array with `interior` in its interior, zero ghosts."""
    nx, ny, nz = interior.shape
    g = np.zeros((nx + 2, ny + 2, nz + 2), order="F")
    g[1:-1, 1:-1, 1:-1] = interior
    return g


def make_solver(case, precision=0):
    """A libfen solver on the GPU configured with `case`: grid, parameters,
    boundary planes and initial fields (ghosts filled by update_ghost_nodes)."""
    from .solver import Solver
    s = Solver(case.dim, case.n, case.L, bc=case.bc, precision=precision)

    def up(fid, arr):
        s.upload(fid, ghosted(arr) if fid <= W else arr)

    case.apply(s, up)
    return s


def load_state(solver, state, dt_o):
    """Uploads interior fields u, v, w, p (ghosts then filled) and dv_o (zero
    unless given), and sets dt_o."""
    from ._lib import DVO_X, DVO_Y, DVO_Z
    for fid in (U, V, W, P):
        solver.upload(fid, ghosted(np.asfortranarray(state[fid])))
        solver.update_ghost_nodes(fid)
    for fid in (DVO_X, DVO_Y, DVO_Z):
        solver.upload(fid, np.asfortranarray(state.get(fid, np.zeros(solver.n))))
    solver.set_dt_o(dt_o)


def interior_velocity(solver):
    return [np.ascontiguousarray(solver.download(f)[1:-1, 1:-1, 1:-1]) for f in (U, V, W)]


def _coords(n, delta):
    return np.arange(1, n + 1, dtype=np.float64)


def tgv2d(N=128, nu=0.01, dt=1e-3, steps=1000):
    """2-D Taylor-Green vortex, doubly periodic [0,2pi]^2.

    u = sin x cos y at (i d, (j-1/2) d), v = -cos x sin y at ((i-1/2) d, j d),
    p = -(cos 2x + cos 2y)/4 at centres (the stated initial data of the
    2-D benchmark case; its pressure sign is washed out by the first projection).
    """
    L = 2.0 * math.pi
    d = L / float(np.float32(N))
    i = np.arange(1, N + 1, dtype=np.float64)
    I, J = np.meshgrid(i, i, indexing="ij")
    u = (np.sin(I * d) * np.cos((J - 0.5) * d))[:, :, None]
    v = (-np.cos((I - 0.5) * d) * np.sin(J * d))[:, :, None]
    p = (-0.25 * (np.cos(2.0 * (I - 0.5) * d) + np.cos(2.0 * (J - 0.5) * d)))[:, :, None]
    return Case("tgv2d", 2, (N, N, 1), (L, L, L / N), [BC_PERIODIC] * 6, 1.0, nu, (0.0, 0.0, 0.0), dt, dt, steps,
                {U: u, V: v, P: p})


def tgv3d(N=256, Re=1600.0, steps=200, eps=1e-3, cfl=0.5, Uref=2.0, seed=SEED):
    """3-D Taylor-Green vortex, triply periodic [0,2pi]^3, plus a seeded
    uniform perturbation eps*U(-1,1) on u, v, w.

    dt follows set_timestep with CFL = 0.5 and U = 2 (dt = delta/4 here).
    Initial pressure p = (cos 2x + cos 2y)(cos 2z + 2)/16 at centres.
    """
    L = 2.0 * math.pi
    d = L / float(np.float32(N))
    nu = 1.0 / Re
    c = (np.arange(1, N + 1, dtype=np.float64) - 0.5) * d
    f = np.arange(1, N + 1, dtype=np.float64) * d
    n3 = N ** 3

    def outer3(a, b, cc):
        return np.asfortranarray((a[:, None, None] * b[None, :, None]) * cc[None, None, :])

    u = outer3(np.sin(f), np.cos(c), np.cos(c))
    v = outer3(-np.cos(c), np.sin(f), np.cos(c))
    w = np.zeros((N, N, N), order="F")
    for comp, arr in enumerate((u, v, w)):
        arr += eps * (2.0 * uniform01(comp, n3, seed).reshape((N, N, N), order="F") - 1.0)
    c2 = np.cos(2.0 * c)
    p = np.asfortranarray(((c2[:, None, None] + c2[None, :, None]) * (c2[None, None, :] + 2.0)) / 16.0)
    dt_conv = cfl * d / Uref
    dt_visc = (1.0 / 6.0) * d * d * 1.0 / nu
    dt = min(dt_conv, dt_visc)
    return Case("tgv3d", 3, (N, N, N), (L, L, L), [BC_PERIODIC] * 6, 1.0, nu, (0.0, 0.0, 0.0), dt, dt, steps,
                {U: np.asfortranarray(u), V: np.asfortranarray(v), W: np.asfortranarray(w), P: np.asfortranarray(p)},
                cfl=cfl)


def isotropic(N=128, nu=1.0e-2, amp=1.0e-2, seed=SEED):
    """Forced isotropic turbulence start (test/large_test/isotropic_turbulence/
    isotropic.f90:30-58, 86-106): triply periodic [0,2pi]^3, viscosity 1e-2,
    ABC velocity evaluated at cell centres plus amp * U[0,1) noise,
    u = cos y + sin z, v = sin x + cos z, w = cos x + sin y; p = 0.
    The program sets dt with set_timestep(U = 1) at the default CFL = 1 and
    then switches to constant CFL = 0.9 (`cfl_run`)."""
    L = 2.0 * math.pi
    d = L / float(np.float32(N))
    c = (np.arange(1, N + 1, dtype=np.float64) - 0.5) * d
    shp = (N, N, N)
    n3 = N ** 3
    X, Y, Z = c[:, None, None], c[None, :, None], c[None, None, :]
    u = np.broadcast_to(np.cos(Y) + np.sin(Z), shp).copy(order="F")
    v = np.broadcast_to(np.sin(X) + np.cos(Z), shp).copy(order="F")
    w = np.broadcast_to(np.cos(X) + np.sin(Y), shp).copy(order="F")
    for comp, arr in enumerate((u, v, w)):
        arr += amp * uniform01(comp, n3, seed).reshape(shp, order="F")
    p = np.zeros(shp, order="F")
    case = Case("isotropic", 3, shp, (L, L, L), [BC_PERIODIC] * 6, 1.0, nu, (0.0, 0.0, 0.0), 0.0, 0.0, 0,
                {U: u, V: v, W: w, P: p}, cfl=1.0)
    case.set_timestep_U = 1.0
    case.cfl_run = 0.9
    return case


def isotropic_forcing(u, v, w):
    """The program's linear forcing S = 0.1 (u - mean u), per component, from
    the interior velocity before each step (isotropic.f90:115-150)."""
    n = u.size
    return [0.1 * (f - f.sum() / n) for f in (u, v, w)]


def channel(nx=512, ny=512, nz=256, Ub=15.7, noise=0.05, nu=1.0 / 180.0, steps=100, cfl=0.5, seed=SEED):
    """Plane channel in FEN axis order: x, y periodic, walls at z (ppn).

    delta = 1/128 at the default size (Lz = 2, h = 1), constant pressure-gradient
    forcing g = (1, 0, 0) (u_tau = 1).  Initial u = 1.5 Ub (1 - (z-1)^2) plus
    Gaussian noise of `noise*Ub` on u, v, w.  dt from set_timestep with
    CFL = 0.5 and U = the centreline velocity 1.5 Ub.
    """
    d = 2.0 / nz
    Lx, Ly, Lz = nx * d, ny * d, 2.0
    zc = (np.arange(1, nz + 1, dtype=np.float64) - 0.5) * d
    n3 = nx * ny * nz
    shp = (nx, ny, nz)
    u = np.broadcast_to(1.5 * Ub * (1.0 - (zc - 1.0) ** 2), shp).copy(order="F")
    v = np.zeros(shp, order="F")
    w = np.zeros(shp, order="F")
    for comp, arr in enumerate((u, v, w)):
        arr += noise * Ub * gaussian(comp, n3, seed).reshape(shp, order="F")
    w[:, :, -1] = 0.0   # the top face of the last cell is the wall
    p = np.zeros(shp, order="F")
    dt_conv = cfl * d / (1.5 * Ub)
    dt_visc = (1.0 / 6.0) * d * d / nu
    dt = min(dt_conv, dt_visc)
    bc = [BC_PERIODIC, BC_PERIODIC, BC_PERIODIC, BC_PERIODIC, BC_WALL, BC_WALL]
    return Case("channel", 3, shp, (Lx, Ly, Lz), bc, 1.0, nu, (1.0, 0.0, 0.0), dt, dt, steps,
                {U: u, V: v, W: w, P: p}, cfl=cfl)


def _noise(shp, amp, seed, first=0):
    n = int(np.prod(shp))
    return [amp * gaussian(first + c, n, seed).reshape(shp, order="F") for c in range(3)]


def _dt(d, Uref, nu, cfl):
    """Convective limit, and a viscous one inside the stability region of
    explicit AB2 diffusion in 3-D (nu dt / d^2 <= 1/12)."""
    return min(cfl * d / Uref, 0.05 * d * d / nu)


def cavity2d(nx=64, ny=64, Re=100.0, steps=50, noise=0.01, cfl=0.5, seed=SEED):
    """Lid-driven cavity, walls on the four sides (nn): the top wall moves
    with u = 1.  Initial u, v: seeded noise, zero on the wall faces."""
    d = 1.0 / nx
    shp = (nx, ny, 1)
    u, v, _ = _noise(shp, noise, seed)
    u[-1, :, :] = 0.0   # the right face of the last cell is the wall
    v[:, -1, :] = 0.0
    p = np.zeros(shp, order="F")
    lid = np.ones((nx + 2, 3))
    bc = [BC_WALL] * 4 + [BC_PERIODIC] * 2
    return Case("cavity2d", 2, shp, (1.0, ny * d, d), bc, 1.0, 1.0 / Re, (0.0, 0.0, 0.0), _dt(d, 1.0, 1.0 / Re, cfl),
                _dt(d, 1.0, 1.0 / Re, cfl), steps, {U: u, V: v, P: p}, cfl=cfl, planes={(U, 3): lid})


def channel2d(nx=128, ny=64, Ub=1.0, Re=100.0, steps=50, noise=0.05, cfl=0.5, inflow=False, seed=SEED):
    """2-D channel, x periodic (pn).  Walls at y with the forcing g = (1, 0);
    or, with `inflow`, a parabolic v profile entering at the bottom face and
    leaving through an outflow face at the top (FEN's Inflow/Outflow pair)."""
    d = 1.0 / ny
    shp = (nx, ny, 1)
    yc = (np.arange(1, ny + 1, dtype=np.float64) - 0.5) * d
    xc = (np.arange(1, nx + 1, dtype=np.float64) - 0.5) * d
    u, v, _ = _noise(shp, noise * Ub, seed)
    if inflow:
        prof = 1.5 * Ub * 4.0 * (xc * (nx * d - xc)) / (nx * d) ** 2
        v += prof[:, None, None]
        plane = np.zeros((nx + 2, 3))
        plane[1:-1, :] = prof[:, None]
        bc = [BC_PERIODIC, BC_PERIODIC, BC_INFLOW, BC_OUTFLOW, BC_PERIODIC, BC_PERIODIC]
        planes, g = {(V, 2): plane}, (0.0, 0.0, 0.0)
    else:
        u += (1.5 * Ub * 4.0 * yc * (1.0 - yc))[None, :, None]
        v[:, -1, :] = 0.0
        bc = [BC_PERIODIC, BC_PERIODIC, BC_WALL, BC_WALL, BC_PERIODIC, BC_PERIODIC]
        planes, g = {}, (1.0, 0.0, 0.0)
    p = np.zeros(shp, order="F")
    nu = Ub / Re
    dt = _dt(d, 1.5 * Ub, nu, cfl)
    return Case("channel2d_io" if inflow else "channel2d", 2, shp, (nx * d, 1.0, d), bc, 1.0, nu, g, dt, dt, steps,
                {U: u, V: v, P: p}, cfl=cfl, planes=planes)


def duct3d(nx=64, ny=32, nz=32, Ub=1.0, Re=100.0, steps=30, noise=0.05, cfl=0.5, z_io=False, seed=SEED):
    """3-D duct, walls at x, y periodic, walls at z (npn): flow along y driven
    by the forcing g = (0, 1, 0) from a seeded perturbation of the laminar
    profile.  With `z_io` the z faces are FEN's Inflow/Outflow pair instead:
    a parabolic w profile enters at the bottom and leaves at the top.  (x
    must stay a wall pair: npn's cosine transform in x fixes a zero normal
    gradient of phi there, which an Outflow face, phi = 0, contradicts.)"""
    d = 1.0 / nz
    shp = (nx, ny, nz)
    zc = (np.arange(1, nz + 1, dtype=np.float64) - 0.5) * d
    xc = (np.arange(1, nx + 1, dtype=np.float64) - 0.5) * d
    Lx = nx * d
    px = 4.0 * xc * (Lx - xc) / Lx ** 2
    u, v, w = _noise(shp, noise * Ub, seed)
    u[-1, :, :] = 0.0
    gy = ny + 2
    if z_io:
        prof = 1.5 * Ub * px
        w += prof[:, None, None]
        plane = np.zeros((nx + 2, gy))
        plane[1:-1, :] = prof[:, None]
        bc = [BC_WALL, BC_WALL, BC_PERIODIC, BC_PERIODIC, BC_INFLOW, BC_OUTFLOW]
        planes, g = {(W, 4): plane}, (0.0, 0.0, 0.0)
    else:
        v += (2.0 * Ub * px[:, None] * (4.0 * zc * (1.0 - zc))[None, :])[:, None, :]
        w[:, :, -1] = 0.0
        bc = [BC_WALL, BC_WALL, BC_PERIODIC, BC_PERIODIC, BC_WALL, BC_WALL]
        planes, g = {}, (0.0, 1.0, 0.0)
    p = np.zeros(shp, order="F")
    nu = Ub / Re
    dt = _dt(d, 2.0 * Ub, nu, cfl)
    return Case("duct3d_zio" if z_io else "duct3d", 3, shp, (Lx, ny * d, 1.0), bc, 1.0, nu, g, dt,
                dt, steps, {U: np.asfortranarray(u), V: np.asfortranarray(v), W: np.asfortranarray(w), P: p},
                cfl=cfl, planes=planes)


def cavity3d(nx=32, ny=32, nz=32, Re=100.0, steps=30, noise=0.01, cfl=0.5, seed=SEED):
    """Lid-driven cubic cavity, walls on all six faces (nnn); the top z wall
    moves with u = 1."""
    d = 1.0 / nx
    shp = (nx, ny, nz)
    u, v, w = _noise(shp, noise, seed)
    u[-1, :, :] = 0.0
    v[:, -1, :] = 0.0
    w[:, :, -1] = 0.0
    p = np.zeros(shp, order="F")
    lid = np.ones((nx + 2, ny + 2))
    nu = 1.0 / Re
    dt = _dt(d, 1.0, nu, cfl)
    return Case("cavity3d", 3, shp, (1.0, ny * d, nz * d), [BC_WALL] * 6, 1.0, nu, (0.0, 0.0, 0.0), dt, dt, steps,
                {U: np.asfortranarray(u), V: np.asfortranarray(v), W: np.asfortranarray(w), P: p}, cfl=cfl,
                planes={(U, 5): lid})


def vof_from_distance(dist, nx, ny, delta, beta=1.0):
    """Cell averages of 0.5 (1 + tanh(beta d / delta)) by 2x2 Gauss points, as
    get_vof_from_distance (reference src/volume_of_fluid.f90:665-706)."""
    rp, rm = 0.5 * (1.0 + 1.0 / math.sqrt(3.0)), 0.5 * (1.0 - 1.0 / math.sqrt(3.0))
    X, Y = np.meshgrid((np.arange(1, nx + 1) - 0.5) * delta, (np.arange(1, ny + 1) - 0.5) * delta, indexing="ij")
    xm, xp = X + delta * (rm - 0.5), X + delta * (rp - 0.5)
    ym, yp = Y + delta * (rm - 0.5), Y + delta * (rp - 0.5)
    t = lambda x, y: 0.5 * (1.0 + np.tanh(beta * dist(x, y) / delta))
    return 0.5 * (0.5 * (t(xm, ym) + t(xp, ym)) + 0.5 * (t(xm, yp) + t(xp, yp)))


def rayleigh_taylor(nx=128, ny=512, steps=100):
    """The reference's Rayleigh-Taylor test (test/small_test/multiphase/
    Rayleigh-Taylor): Lx = 1, Ly = 4, x periodic, walls at y, air-like
    rho_0 = 0.1694 below the interface y = Ly/2 + 0.05 cos(2 pi x), rho_1 =
    1.225 above, mu = 3.13e-3 in both, g = -9.80665 in y, no surface tension.
    Returns (case, multiphase properties, interior vof)."""
    Lx, Ly = 1.0, 4.0
    d = Lx / nx
    vof = vof_from_distance(lambda x, y: y - 0.05 * np.cos(2.0 * math.pi * x / Lx) - Ly / 2.0, nx, ny, d)
    shp = (nx, ny, 1)
    z = np.zeros(shp, order="F")
    bc = [BC_PERIODIC, BC_PERIODIC, BC_WALL, BC_WALL, BC_PERIODIC, BC_PERIODIC]
    props = dict(rho0=0.1694, rho1=1.225, mu0=3.13e-3, mu1=3.13e-3, sigma=0.0)
    case = Case("rayleigh_taylor", 2, shp, (Lx, Ly, d), bc, 1.0, 3.13e-3, (0.0, -9.80665, 0.0), 0.0, 0.0, steps,
                {U: z, V: z.copy(), P: z.copy()}, cfl=1.0)
    return case, props, vof[:, :, None]


def interior(shape_ghosted):
    return tuple(slice(1, s - 1) for s in shape_ghosted)

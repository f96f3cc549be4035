"""Pinning the CPU oracle (oracle/fen_oracle.c) without the reference binary.

The reference cannot be built here (its Poisson and halo modules need FFTW3
and 2decomp&FFT, which are not installed), and it ships no golden vectors
for this path.  Each oracle function is therefore checked against something
that fixes it independently:

* FFTs: brute-force DFT in long double, FFTW's documented conventions.
* Thomas sweep: dense linear solve; the singular (0,0) mode gives last pivot 0.
* Poisson solvers: the discrete-Laplacian inverse (apply FEN's 7-point
  Laplacian with the solver's boundary rules to the solution and get the
  right-hand side back, minus its mean), and the reference's own
  convergence-rate test (order > 1.8,
  test/small_test/poisson/convergence_rate/postpro.py:28).
* Projection: test/small_test/poisson/projection/projection.f90 (max div <= 1e-11).
* Advection operator: closed form of test/small_test/navier_stokes/advection.
* Full step: 2-D Taylor-Green decay (order > 1.8,
  test/small_test/navier_stokes/taylor_green_vortex/postpro.py:62) and the
  3-D ABC flow (order > 1.5, test/large_test/ABC/postpro.py:42).
* Ghost rules: what each boundary condition means at the face.
* Restart: 10 steps == 5 steps + state reload + 5 steps, bitwise
  (test/small_test/io/run.sh:6-16).
"""
import math

import numpy as np
import pytest

from fen_amd import cases
from fen_amd._lib import BC_PERIODIC, BC_WALL, DVO_X, DVO_Y, DVO_Z, P, PHI, U, V, W
from oracle import oracle as O
from tests.common import make_oracle


# ----------------------------------------------------------------- FFT
def _dft_longdouble(x, sign):
    n = len(x)
    k = np.arange(n, dtype=np.longdouble)
    pi = np.longdouble("3.14159265358979323846264338327950288")
    ang = sign * 2 * pi * (np.outer(k, k) % n) / n
    xr = np.real(x).astype(np.longdouble)
    xi = np.imag(x).astype(np.longdouble)
    c, s = np.cos(ang), np.sin(ang)
    return (c @ xr - s @ xi).astype(np.float64) + 1j * (s @ xr + c @ xi).astype(np.float64)


@pytest.mark.parametrize("n", [8, 16, 32, 64, 128])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_c2c_against_bruteforce_dft(n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = O.fft_c2c(x, sign)
    ref = _dft_longdouble(x, sign)
    assert np.abs(y - ref).max() <= 4e-16 * np.abs(x).sum()


@pytest.mark.parametrize("n", [8, 16, 64, 256])
def test_fft_r2c_c2r_conventions(n):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(n)
    X = O.fft_r2c(x)
    assert X.shape == (n // 2 + 1,)
    ref = _dft_longdouble(x.astype(complex), -1)[: n // 2 + 1]
    assert np.abs(X - ref).max() <= 4e-16 * np.abs(x).sum()
    # c2r is the unnormalised inverse: c2r(r2c(x)) = n x
    y = O.fft_c2r(X, n)
    assert np.abs(y - n * x).max() <= 1e-13 * n * np.abs(x).max()
    # FFTW's c2r ignores the imaginary parts of bins 0 and n/2
    X2 = X.copy()
    X2[0] += 3.0j
    X2[-1] -= 2.0j
    assert np.array_equal(O.fft_c2r(X2, n), y)


@pytest.mark.parametrize("n", [256, 512])
def test_fft_parseval_and_roundtrip(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = O.fft_c2c(x, -1)
    assert abs(np.sum(np.abs(X) ** 2) / n - np.sum(np.abs(x) ** 2)) < 1e-12 * np.sum(np.abs(x) ** 2)
    back = O.fft_c2c(X, 1) / n
    assert np.abs(back - x).max() < 1e-14 * 10


# ------------------------------------------------------------- Thomas
def _tridiag(n, delta):
    d2 = delta * delta
    a = np.full(n, 1.0 / d2)
    b = np.full(n, -(2.0 / d2))
    c = np.full(n, 1.0 / d2)
    b[0] += a[0]
    b[-1] += c[-1]
    a[0] = 0.0
    c[-1] = 0.0
    return a, b, c


@pytest.mark.parametrize("n", [2, 5, 16, 64])
def test_thomas_against_dense_solve(n):
    rng = np.random.default_rng(n)
    delta = 1.0 / n
    a, b, c = _tridiag(n, delta)
    lam_x, lam_y = -3.7 / delta, -1.1
    r = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x = O.thomas(a, b, c, lam_x, lam_y, r)
    M = np.diag(b + lam_x + lam_y) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    ref = np.linalg.solve(M, r)
    assert np.abs(x - ref).max() < 1e-12 * np.abs(ref).max()


def test_thomas_singular_mode_has_zero_last_pivot():
    # lambda = 0 with Neumann ends: the matrix is singular and FEN's sweep hits
    # an exactly zero last pivot (src/poisson.f90:1114-1119) -> x_n = 0.
    n = 32
    a, b, c = _tridiag(n, 1.0 / n)
    r = np.random.default_rng(1).standard_normal(n) + 0j
    r -= r.mean()   # compatible right-hand side
    x = O.thomas(a, b, c, 0.0, 0.0, r)
    assert x[-1] == 0.0
    M = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    assert np.abs(M @ x - r).max() < 1e-10 * np.abs(r).max() * n * n


# ------------------------------------------------- Poisson solvers
def _laplacian(phi, delta, z_neumann=False, dim=3):
    """FEN's 7-point Laplacian (src/fields.f90:256-294) with periodic x, y and
    periodic or Neumann (ghost = first cell) z."""
    def pad(a, axis, neumann):
        if neumann:
            lo = np.take(a, [0], axis=axis)
            hi = np.take(a, [-1], axis=axis)
        else:
            lo = np.take(a, [-1], axis=axis)
            hi = np.take(a, [0], axis=axis)
        return np.concatenate([lo, a, hi], axis=axis)

    out = np.zeros_like(phi)
    for ax in range(dim):
        g = pad(phi, ax, z_neumann and ax == 2)
        sl = [slice(None)] * 3
        sl[ax] = slice(2, None)
        up = g[tuple(sl)]
        sl[ax] = slice(0, -2)
        dn = g[tuple(sl)]
        out += (up - 2.0 * phi + dn)
    return out / (delta * delta)


@pytest.mark.parametrize("kind,dim,n,bc", [
    ("pp", 2, (32, 16, 1), [BC_PERIODIC] * 6),
    ("ppp", 3, (16, 8, 32), [BC_PERIODIC] * 6),
    ("ppn", 3, (16, 8, 12), [BC_PERIODIC] * 4 + [BC_WALL] * 2),
])
def test_poisson_inverts_fen_laplacian(kind, dim, n, bc):
    rng = np.random.default_rng(2)
    L = (1.0, n[1] / n[0], n[2] / n[0])
    o = O.Oracle(dim, n, L, bc=bc)
    assert o.solver == kind
    f = np.asfortranarray(rng.standard_normal(o.n))
    if kind == "ppn":
        # the (0,0) z-column is singular (Neumann ends): FEN's sweep solves it
        # only for a compatible, zero-mean right-hand side
        f -= f.mean()
    phi = o.poisson_solve(f.copy(order="F"), 0)
    lap = _laplacian(phi, o.delta, z_neumann=(kind == "ppn"), dim=dim)
    target = f - f.mean()
    assert np.abs(lap - target).max() < 1e-11 * np.abs(target).max()
    if kind in ("pp", "ppn"):   # these solvers remove the mean of phi
        assert abs(phi.mean()) < 1e-12 * np.abs(phi).max()


def _laplacian_axes(phi, delta, neumann):
    """FEN's Laplacian with, per axis, periodic wrap or Neumann ends (ghost =
    first/last cell): the operator the Neumann solvers invert (tridiagonal
    b1 += a1, bn += cn, src/poisson.f90:214-226; DCT-II modes, :265)."""
    out = np.zeros_like(phi)
    for ax, neu in enumerate(neumann):
        if phi.shape[ax] == 1:
            continue
        lo = np.take(phi, [0] if neu else [-1], axis=ax)
        hi = np.take(phi, [-1] if neu else [0], axis=ax)
        g = np.concatenate([lo, phi, hi], axis=ax)
        sl = [slice(None)] * 3
        sl[ax] = slice(2, None)
        up = g[tuple(sl)]
        sl[ax] = slice(0, -2)
        out += (up - 2.0 * phi + g[tuple(sl)])
    return out / (delta * delta)


@pytest.mark.parametrize("kind,dim,n,bc", [
    ("pn", 2, (32, 16, 1), [BC_PERIODIC] * 2 + [BC_WALL] * 2 + [BC_PERIODIC] * 2),
    ("nn", 2, (16, 32, 1), [BC_WALL] * 4 + [BC_PERIODIC] * 2),
    ("npn", 3, (16, 8, 12), [BC_WALL] * 2 + [BC_PERIODIC] * 2 + [BC_WALL] * 2),
    ("nnn", 3, (8, 16, 12), [BC_WALL] * 6),
])
def test_neumann_poisson_solvers_invert_fen_laplacian(kind, dim, n, bc):
    """pn, nn, npn, nnn (src/poisson.f90:306-596, 1177-1451): DCT-II/III along
    Neumann axes, FFT along periodic ones, Thomas along the last; the result
    satisfies FEN's Laplacian with Neumann ends to round-off for a zero-mean
    right-hand side (the constant mode is singular)."""
    rng = np.random.default_rng(9)
    L = (1.0, n[1] / n[0], n[2] / n[0])
    o = O.Oracle(dim, n, L, bc=bc)
    assert o.solver == kind
    f = np.asfortranarray(rng.standard_normal(o.n))
    f -= f.mean()
    phi = o.poisson_solve(f.copy(order="F"), 0)
    neumann = [kind[0] == "n", kind[1] == "n", dim == 3 and kind[2] == "n"]
    lap = _laplacian_axes(phi, o.delta, neumann)
    assert np.abs(lap - f).max() < 1e-10 * np.abs(f).max(), np.abs(lap - f).max()
    if kind == "pn":   # pn removes the interior mean (:408)
        assert abs(phi.mean()) < 1e-12 * np.abs(phi).max()


def test_neumann_solver_dct_conventions():
    """The x transform pair of nn is FFTW's REDFT10/REDFT01: a single cosine
    mode cos(pi k (i - 1/2)/n) is an eigenvector of the Neumann Laplacian with
    eigenvalue 2(cos(pi k/n) - 1)/delta^2, and the solve divides by it."""
    n = (32, 16, 1)
    o = O.Oracle(2, n, (1.0, 0.5, 1.0), bc=[BC_WALL] * 4 + [BC_PERIODIC] * 2)
    i = np.arange(1, n[0] + 1)[:, None, None]
    j = np.arange(1, n[1] + 1)[None, :, None]
    for kx, ky in ((1, 0), (5, 3), (31, 15)):
        mode = np.cos(np.pi * kx * (i - 0.5) / n[0]) * np.cos(np.pi * ky * (j - 0.5) / n[1]) * np.ones(n)
        lam = 2 * (np.cos(np.pi * kx / n[0]) - 1) / o.delta ** 2 + 2 * (np.cos(np.pi * ky / n[1]) - 1) / o.delta ** 2
        phi = o.poisson_solve(np.asfortranarray(mode), 0)
        assert np.abs(phi - mode / lam).max() < 1e-11 * np.abs(mode / lam).max(), (kx, ky)


@pytest.mark.parametrize("case", ["ppp", "ppn", "pp"])
def test_poisson_convergence_rate(case):
    """test/small_test/poisson/convergence_rate/convergence_rate.f90 (order > 1.8)."""
    Ns, errs = [16, 32, 64], []
    tp = 2.0 * math.pi
    for N in Ns:
        if case == "pp":
            o = O.Oracle(2, (N, N, 1), (1.0, 1.0, 1.0 / N))
        elif case == "ppp":
            o = O.Oracle(3, (N, N, N), (1.0, 1.0, 1.0))
        else:
            o = O.Oracle(3, (N, N, N), (1.0, 1.0, 1.0), bc=[0, 0, 0, 0, 1, 1])
        d = o.delta
        x = (np.arange(1, N + 1) - 0.5) * d
        if case == "pp":
            X, Y = np.meshgrid(x, x, indexing="ij")
            rhs = -8 * math.pi ** 2 * np.sin(tp * X) * np.cos(tp * Y)
            sol = np.sin(tp * X) * np.cos(tp * Y)
            rhs, sol = rhs[:, :, None], sol[:, :, None]
        else:
            X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
            if case == "ppp":
                rhs = -12 * math.pi ** 2 * np.sin(tp * X) * np.cos(tp * Y) * np.sin(tp * Z)
                sol = np.sin(tp * X) * np.cos(tp * Y) * np.sin(tp * Z)
            else:
                rhs = -4 * math.pi ** 2 * (np.sin(tp * X) + np.sin(tp * Y) + np.cos(tp * Z))
                sol = np.sin(tp * X) + np.sin(tp * Y) + np.cos(tp * Z)
        phi = o.poisson_solve(np.asfortranarray(rhs), 0)
        errs.append(np.abs(phi - sol).max())
    order = -np.polyfit(np.log(Ns), np.log(errs), 1)[0]
    assert order > 1.8, (case, errs)


@pytest.mark.parametrize("walls", [False, True])
def test_projection_divergence_free(walls):
    """test/small_test/poisson/projection/projection.f90: random u, phi = -div u,
    solve, u += grad phi  ->  max |div u| <= 1e-11."""
    N = 32
    bc = [0, 0, 0, 0, 1, 1] if walls else [0] * 6
    o = O.Oracle(3, (N, N, N), (1.0, 1.0, 1.0), bc=bc)
    rng = np.random.default_rng(9)
    worst = 0.0
    for trial in range(5):
        for fid in (U, V, W):
            o.field(fid)[1:-1, 1:-1, 1:-1] = rng.random((N, N, N))
            o.update_ghost_nodes(fid)
        u, v, w = o.field(U), o.field(V), o.field(W)
        div = o.divergence(u, v, w)
        phi = o.field(PHI)
        phi[1:-1, 1:-1, 1:-1] = -div
        o.poisson_solve(phi, 1)
        o.update_ghost_nodes(PHI)
        gx, gy, gz = o.gradient(phi)
        u[1:-1, 1:-1, 1:-1] += gx
        v[1:-1, 1:-1, 1:-1] += gy
        w[1:-1, 1:-1, 1:-1] += gz
        for fid in (U, V, W):
            o.update_ghost_nodes(fid)
        worst = max(worst, np.abs(o.divergence(u, v, w)).max())
    assert worst <= 1e-11


# ------------------------------------------------------ operators
def test_advection_closed_form_order():
    """test/small_test/navier_stokes/advection/advection.f90: for u = sin x cos y,
    v = -cos x sin y, add_advection gives -(u.grad)u to second order."""
    errs, Ns = [], [16, 32, 64, 128]
    for N in Ns:
        L = 2 * math.pi
        o = O.Oracle(2, (N, N, 1), (L, L, L / N))
        d = o.delta
        i = np.arange(1, N + 1)
        I, J = np.meshgrid(i, i, indexing="ij")
        o.field(U)[1:-1, 1:-1, 1] = np.sin(I * d) * np.cos((J - 0.5) * d)
        o.field(V)[1:-1, 1:-1, 1] = -np.cos((I - 0.5) * d) * np.sin(J * d)
        o.update_ghost_nodes(U)
        o.update_ghost_nodes(V)
        rx, ry, _ = o.add_advection()
        x, y = I * d, (J - 0.5) * d
        advu = np.sin(x) * np.cos(y) * np.cos(x) * np.cos(y) + (-np.cos(x) * np.sin(y)) * (-np.sin(x) * np.sin(y))
        x, y = (I - 0.5) * d, J * d
        advv = np.sin(x) * np.cos(y) * np.sin(x) * np.sin(y) + (-np.cos(x) * np.sin(y)) * (-np.cos(x) * np.cos(y))
        errs.append(max(np.abs(rx[:, :, 0] + advu).max(), np.abs(ry[:, :, 0] + advv).max()))
    order = -np.polyfit(np.log(Ns), np.log(errs), 1)[0]
    assert order > 1.8, errs


def test_advection_3d_closed_form_order():
    """add_advection in 3-D (src/navier_stokes.f90:297-347) against the continuum
    divergence form d(u u_j)/dx_j of a field that is neither solenoidal nor
    Beltrami, so no part of the term can hide in a gradient: each component
    converges at second order to its own closed form (derivatives by sympy)."""
    import sympy as sp
    X, Y, Z = sp.symbols("x y z")
    ue = sp.sin(X + 0.3) * sp.cos(2 * Y) * sp.cos(Z)
    ve = sp.cos(X) * sp.sin(Y + 0.7) * sp.sin(2 * Z) + 0.5
    we = sp.sin(2 * X) * sp.cos(Y) * sp.sin(Z + 0.2) - 0.25
    vel = (ue, ve, we)
    adv = [sp.lambdify((X, Y, Z), sum(sp.diff(vel[a] * vel[b], s) for b, s in enumerate((X, Y, Z))), "numpy")
           for a in range(3)]
    fun = [sp.lambdify((X, Y, Z), e, "numpy") for e in vel]
    errs, Ns = [], [16, 32, 64]
    for N in Ns:
        L = 2 * math.pi
        o = O.Oracle(3, (N, N, N), (L, L, L))
        d = o.delta
        f = np.arange(1, N + 1) * d           # face coordinate of index i
        c = (np.arange(1, N + 1) - 0.5) * d   # centre coordinate
        locs = [(f, c, c), (c, f, c), (c, c, f)]
        for a, fid in enumerate((U, V, W)):
            gx, gy, gz = np.meshgrid(*locs[a], indexing="ij")
            o.field(fid)[1:-1, 1:-1, 1:-1] = fun[a](gx, gy, gz)
            o.update_ghost_nodes(fid)
        r = o.add_advection()
        e = []
        for a in range(3):
            gx, gy, gz = np.meshgrid(*locs[a], indexing="ij")
            e.append(np.abs(r[a] + adv[a](gx, gy, gz)).max())
        errs.append(e)
    errs = np.array(errs)
    for a in range(3):
        order = -np.polyfit(np.log(Ns), np.log(errs[:, a]), 1)[0]
        assert order > 1.8, (a, errs[:, a])


def test_ghost_rules_mean_what_they_say():
    """Dirichlet tangential / centre: the face value (ghost+first)/2 is bc;
    Dirichlet normal: the face itself is bc (high side overwrites f(N));
    Neumann: ghost equals the first cell; periodic: ghost equals the far cell."""
    rng = np.random.default_rng(4)
    n = (8, 6, 10)
    o = O.Oracle(3, n, (1.0, 0.75, 1.25), bc=[0, 0, 0, 0, 1, 1])   # walls in z
    for fid in (U, W, P):
        o.field(fid)[1:-1, 1:-1, 1:-1] = rng.standard_normal(n)
        for face in range(6):
            o.bc_plane(fid, face)[...] = rng.standard_normal(o.bc_plane(fid, face).shape)
    for fid in (U, W, P):
        o.update_ghost_nodes(fid)
    u, w, p = o.field(U), o.field(W), o.field(P)
    bcF, bcB = o.bc_plane(U, 4), o.bc_plane(U, 5)
    # u is tangential to the z walls: (ghost + first)/2 == bc on the face
    assert np.allclose(0.5 * (u[:, :, 0] + u[:, :, 1]), bcF, rtol=0, atol=1e-14)
    assert np.allclose(0.5 * (u[:, :, -1] + u[:, :, -2]), bcB, rtol=0, atol=1e-14)
    # w is normal: w(0) = bc front, w(N) = w(N+1) = bc back
    assert np.array_equal(w[:, :, 0], o.bc_plane(W, 4))
    assert np.array_equal(w[:, :, -2], o.bc_plane(W, 5))
    assert np.array_equal(w[:, :, -1], o.bc_plane(W, 5))
    # p is Neumann at walls
    assert np.array_equal(p[:, :, 0], p[:, :, 1]) and np.array_equal(p[:, :, -1], p[:, :, -2])
    # periodic x and y, including edges and corners
    assert np.array_equal(p[0, 1:-1, 1:-1], p[-2, 1:-1, 1:-1]) and np.array_equal(p[-1, 1:-1, 1:-1], p[1, 1:-1, 1:-1])
    assert np.array_equal(u[0, 0, 1:-1], u[-2, -2, 1:-1])


# ------------------------------------------------------- full step
def test_taylor_green_2d_convergence():
    """taylor_green_vortex.f90 (nu = 1, t = 0.3): order > 1.8 against exp(-2 nu t)."""
    errs, Ns = [], [16, 32, 64]
    for N in Ns:
        L = 2 * math.pi
        o = O.Oracle(2, (N, N, 1), (L, L, L / N))
        o.set_params(viscosity=1.0)
        d = o.delta
        i = np.arange(1, N + 1)
        I, J = np.meshgrid(i, i, indexing="ij")
        o.field(U)[1:-1, 1:-1, 1] = -np.cos(I * d) * np.sin((J - 0.5) * d)
        o.field(V)[1:-1, 1:-1, 1] = np.sin((I - 0.5) * d) * np.cos(J * d)
        o.field(P)[1:-1, 1:-1, 1] = -0.25 * (np.cos(2 * (I - 0.5) * d) + np.cos(2 * (J - 0.5) * d))
        for f in (P, U, V):
            o.update_ghost_nodes(f)
        dt = o.set_timestep(2.0)
        t, s = 0.0, 0
        while t < 0.3:
            s += 1
            t += dt
            dt, md, mc = o.step(s, dt)
            assert abs(md) < 1e-12
        sol = -np.cos(I * d) * np.sin((J - 0.5) * d) * math.exp(-2 * t)
        errs.append(np.abs(o.field(U)[1:-1, 1:-1, 1] - sol).max())
    order = -np.polyfit(np.log(Ns), np.log(errs), 1)[0]
    assert order > 1.8, errs


def test_abc_flow_3d_convergence():
    """test/large_test/ABC/ABC.f90 (nu = 0.1, t = 0.1, dt = set_timestep/2): order > 1.5."""
    errs, Ns = [], [16, 32]
    for N in Ns:
        L = 2 * math.pi
        o = O.Oracle(3, (N, N, N), (L, L, L))
        o.set_params(viscosity=0.1)
        d = o.delta
        c = (np.arange(0, N + 2) - 0.5) * d
        X, Y, Z = np.meshgrid(c[1:-1], c[1:-1], c[1:-1], indexing="ij")
        o.field(U)[1:-1, 1:-1, 1:-1] = np.sin(Z) + np.cos(Y)
        o.field(V)[1:-1, 1:-1, 1:-1] = np.sin(X) + np.cos(Z)
        o.field(W)[1:-1, 1:-1, 1:-1] = np.cos(X) + np.sin(Y)
        o.field(P)[1:-1, 1:-1, 1:-1] = -(np.cos(X) * np.sin(Y) + np.sin(X) * np.cos(Z) + np.cos(Y) * np.sin(Z))
        dt = o.set_timestep(2.0) / 2.0
        for f in (U, V, W, P):
            o.update_ghost_nodes(f)
        t, s = 0.0, 0
        while t <= 0.1:
            s += 1
            t += dt
            o.step(s, dt)
        sol = (np.cos(Y) + np.sin(Z)) * math.exp(-0.1 * 0.1)
        errs.append(np.abs(o.field(U)[1:-1, 1:-1, 1:-1] - sol).max())
    order = -np.polyfit(np.log(Ns), np.log(errs), 1)[0]
    assert order > 1.5, errs


def test_restart_is_bitwise():
    """io/run.sh: 10 steps straight == 5 steps, reload (p, u, v, w, dv_o), 5 steps."""
    case = cases.tgv3d(16, steps=10)
    a = make_oracle(case)
    for s in range(1, 11):
        a.step(s, case.dt)
    b = make_oracle(case)
    for s in range(1, 6):
        b.step(s, case.dt)
    saved = {f: b.field(f)[1:-1, 1:-1, 1:-1].copy() if f <= W else b.field(f).copy()
             for f in (P, U, V, W, DVO_X, DVO_Y, DVO_Z)}
    c = make_oracle(case)
    for f, arr in saved.items():
        if f <= W:
            c.field(f)[1:-1, 1:-1, 1:-1] = arr
        else:
            c.field(f)[...] = arr
    for f in (P, U, V, W):
        c.update_ghost_nodes(f)
    for s in range(6, 11):
        c.step(s, case.dt)
    for f in (P, U, V, W):
        assert np.array_equal(c.field(f), a.field(f))


# --------------------------------------------------------- fixtures
@pytest.mark.parametrize("name", ["tgv2d_16_10steps", "tgv3d_16_5steps", "channel_16x8x8_5steps"])
def test_oracle_reproduces_golden_fixtures(name):
    import os
    from tests.golden.make_golden import run
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", name + ".npz"))
    got = run(name)
    for k in ref.files:
        assert np.array_equal(got[k], ref[k]), k


def test_splitmix64_matches_the_published_algorithm():
    def sm(x):
        m = (1 << 64) - 1
        z = x & m
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        return z ^ (z >> 31)
    g = 0x9E3779B97F4A7C15
    seed = cases.SEED
    u = cases.uniform01(2, 5)
    for q in range(5):
        ctr = (seed + (2 * (1 << 40) + q + 1) * g) & ((1 << 64) - 1)
        assert u[q] == (sm(ctr) >> 11) * 2.0 ** -53

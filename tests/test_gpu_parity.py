"""GPU (libfen) against the CPU oracle on identical seeded inputs.

Bitwise where the reference arithmetic has no transform in it (ghost fill,
advection, the momentum right-hand side); within 1e-12 relative for
anything that goes through an FFT (the oracle's radix-2 and the GPU's
radix-8 Stockham transforms round differently), as fixed for this path.
"""
import math

import numpy as np
import pytest

from fen_amd import cases
from fen_amd._lib import BC_PERIODIC, BC_WALL, DVO_X, DVO_Y, DVO_Z, P, PHI, U, V, W
from oracle import oracle as O
from tests.common import ghosted, live, make_gpu, make_oracle, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _random_state(case, rng):
    fields = {}
    for fid in ([U, V, W, P] if case.dim == 3 else [U, V, P]):
        fields[fid] = np.asfortranarray(rng.standard_normal(case.n))
    case.fields = fields
    return case


@pytest.mark.parametrize("dim,bc", [
    (2, [BC_PERIODIC] * 6),
    (3, [BC_PERIODIC] * 6),
    (3, [BC_PERIODIC] * 4 + [BC_WALL] * 2),
])
def test_ghost_fill_bitwise(dim, bc):
    rng = np.random.default_rng(11)
    n = (16, 8, 1) if dim == 2 else (16, 8, 8)
    case = cases.Case("g", dim, n, (1.0, 0.5, 0.5), bc, 1.0, 0.1, (0, 0, 0), 1e-3, 1e-3, 0, {})
    _random_state(case, rng)
    o = make_oracle(case)
    g = make_gpu(case)
    for fid in case.fields:
        if fid in (U, V, W, P):
            # non-zero boundary values exercise the Dirichlet rules
            for face in range(6 if dim == 3 else 4):
                plane = rng.standard_normal(o.bc_plane(fid, face).shape)
                o.bc_plane(fid, face)[...] = plane
                g.set_bc_plane(fid, face, plane)
            o.update_ghost_nodes(fid)
            g.update_ghost_nodes(fid)
        a = live(g.download(fid), dim)
        b = live(o.field(fid), dim)
        assert np.array_equal(a, b), (fid, np.argwhere(a != b)[:5])


@pytest.mark.parametrize("dim", [2, 3])
def test_add_advection_bitwise(dim):
    rng = np.random.default_rng(5)
    n = (32, 16, 1) if dim == 2 else (16, 16, 8)
    case = cases.Case("a", dim, n, (1.0, 0.5, 0.25), [BC_PERIODIC] * 6, 1.0, 0.1, (0, 0, 0), 1e-3, 1e-3, 0, {})
    _random_state(case, rng)
    o = make_oracle(case)
    g = make_gpu(case)
    R0 = [np.asfortranarray(rng.standard_normal(o.n)) for _ in range(3)]
    ro = [r.copy(order="F") for r in R0]
    O.lib().orc_add_advection(o.h, *[O._p(r) for r in ro])
    rg = g.add_advection(R0)
    for d in range(dim):
        assert np.array_equal(rg[d], ro[d]), d


@pytest.mark.parametrize("kind", ["pp", "ppp", "ppn"])
@pytest.mark.parametrize("gl", [0, 1])
def test_poisson_solve_matches_oracle(kind, gl):
    rng = np.random.default_rng(3)
    if kind == "pp":
        dim, n, bc = 2, (64, 32, 1), [BC_PERIODIC] * 6
    elif kind == "ppp":
        dim, n, bc = 3, (32, 16, 64), [BC_PERIODIC] * 6
    else:
        dim, n, bc = 3, (32, 16, 24), [BC_PERIODIC] * 4 + [BC_WALL] * 2
    L = (1.0, n[1] / n[0], n[2] / n[0])
    o = O.Oracle(dim, n, L, bc=bc)
    from fen_amd import Solver
    g = Solver(dim, n, L, bc=bc)
    assert o.solver == kind and g.solver == kind
    shp = (n[0] + 2 * gl, n[1] + 2 * gl, (1 if dim == 2 else n[2]) + 2 * gl)
    f = np.asfortranarray(rng.standard_normal(shp))
    fo = o.poisson_solve(f.copy(order="F"), gl)
    fg = g.solve_poisson(f, gl)
    assert rel_err(fg, fo) < 1e-13


def test_tgv2d_1000_steps_matches_oracle():
    case = cases.tgv2d(128, steps=1000)
    o = make_oracle(case)
    g = make_gpu(case)
    dt = case.dt
    for s in range(1, case.steps + 1):
        dto, mdo, mco = o.step(s, dt)
        dtg, mdg, mcg = g.navier_stokes_solver(s, dt, want_checks=(s % 100 == 0))
        if s % 100 == 0:
            assert abs(mcg - mco) <= 1e-12 * abs(mco)
            assert abs(mdg) < 1e-12 and abs(mdo) < 1e-12
    for fid in (U, V, P):
        e = rel_err(live(g.download(fid), 2), live(o.field(fid), 2))
        assert e < TOL, (fid, e)
    # analytic decay exp(-2 nu t) of the vortex (velocity only)
    t = case.steps * dt
    u = g.download(U)[1:-1, 1:-1, 1]
    d = g.delta
    i = np.arange(1, 129)
    I, J = np.meshgrid(i, i, indexing="ij")
    exact = np.sin(I * d) * np.cos((J - 0.5) * d) * math.exp(-2 * case.viscosity * t)
    assert np.abs(u - exact).max() < 2e-3 * math.exp(-2 * case.viscosity * t)


def test_tgv3d_first_step_rhs_bitwise_and_20_steps():
    case = cases.tgv3d(32, steps=20)
    o = make_oracle(case)
    g = make_gpu(case)
    dt = case.dt
    o.step(1, dt)
    g.navier_stokes_solver(1, dt)
    # dv_o after step 1 is the momentum right-hand side of the initial state: pure stencils
    for fid in (DVO_X, DVO_Y, DVO_Z):
        assert np.array_equal(g.download(fid), o.field(fid)), fid
    for s in range(2, case.steps + 1):
        o.step(s, dt)
        _, mdg, _ = g.navier_stokes_solver(s, dt)
        assert abs(mdg) < 1e-12
    for fid in (U, V, W, P):
        e = rel_err(g.download(fid), o.field(fid))
        assert e < TOL, (fid, e)


def test_channel_small_20_steps():
    case = cases.channel(32, 32, 16, steps=20)
    o = make_oracle(case)
    g = make_gpu(case)
    assert g.solver == "ppn"
    dt = case.dt
    for s in range(1, case.steps + 1):
        o.step(s, dt)
        _, mdg, _ = g.navier_stokes_solver(s, dt)
        assert abs(mdg) < 1e-9
    for fid in (U, V, W, P):
        e = rel_err(g.download(fid), o.field(fid))
        assert e < TOL, (fid, e)


def test_run_to_run_bitwise_256():
    case = cases.tgv3d(256, steps=3)
    a = make_gpu(case)
    a.run(0, 3, case.dt)
    ua = a.download(U)
    a.close()
    b = make_gpu(case)
    b.run(0, 3, case.dt)
    assert np.array_equal(b.download(U), ua)
    md = b.scalars()["maxdiv"]
    assert abs(md) < 1e-10


def _checks_numpy(u, v, w, dim, delta):
    """checks (src/navier_stokes.f90:570-619) evaluated on downloaded ghosted fields."""
    idl = 1.0 / delta
    if dim == 2:
        U, Vv = u[:, :, 1], v[:, :, 1]
        d = (U[1:-1, 1:-1] - U[:-2, 1:-1]) * idl + (Vv[1:-1, 1:-1] - Vv[1:-1, :-2]) * idl
        vel = np.abs(U[1:-1, 1:-1]) + np.abs(Vv[1:-1, 1:-1])
    else:
        I = (slice(1, -1),) * 3
        d = (u[I] - u[:-2, 1:-1, 1:-1]) * idl + (v[I] - v[1:-1, :-2, 1:-1]) * idl
        d = d + (w[I] - w[1:-1, 1:-1, :-2]) * idl
        vel = np.abs(u[I]) + np.abs(v[I]) + np.abs(w[I])
    return d.max(), vel.max()


@pytest.mark.parametrize("which", ["tgv2d", "tgv3d", "channel"])
def test_fused_checks_equal_checks_on_final_fields(which):
    if which == "tgv2d":
        case = cases.tgv2d(64, steps=3)
    elif which == "tgv3d":
        case = cases.tgv3d(32, steps=3)
    else:
        case = cases.channel(32, 16, 16, steps=3)
    g = make_gpu(case)
    for s in range(1, 4):
        dt, md, mc = g.navier_stokes_solver(s, case.dt)
    u, v = g.download(U), g.download(V)
    w = g.download(W) if case.dim == 3 else None
    dmax, vmax = _checks_numpy(u, v, w, case.dim, g.delta)
    assert md == dmax
    assert mc == dt * vmax / g.delta


@pytest.mark.parametrize("name", ["tgv2d_16_10steps", "tgv3d_16_5steps", "channel_16x8x8_5steps"])
def test_gpu_matches_golden_fixtures(name):
    import os
    from tests.golden.make_golden import FIXTURES
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", name + ".npz"))
    make, fields = FIXTURES[name]
    case = make()
    g = make_gpu(case)
    for s in range(1, case.steps + 1):
        dt, md, mc = g.navier_stokes_solver(s, case.dt)
    assert abs(mc - float(ref["maxcfl"])) <= 1e-12 * abs(float(ref["maxcfl"]))
    for f in fields:
        key = {P: "p", U: "u", V: "v", W: "w"}[f]
        e = rel_err(live(g.download(f), case.dim), live(ref[key], case.dim))
        assert e < TOL, (key, e)


@pytest.mark.parametrize("name", ["tgv2d", "tgv3d", "channel"])
def test_step_graph_replay_matches_eager(name, monkeypatch):
    """Steps replayed from the captured HIP graph equal eager steps bit for
    bit, across replays separated by downloads, a dt change (re-capture)
    and a parameter change."""
    case = {"tgv2d": lambda: cases.tgv2d(32, steps=12), "tgv3d": lambda: cases.tgv3d(32, steps=12),
            "channel": lambda: cases.channel(32, 16, 16, steps=12)}[name]()
    comps = [U, V, P] if case.dim == 2 else [U, V, W, P]
    results = []
    for graph in ("0", "1"):
        monkeypatch.setenv("FEN_STEP_GRAPH", graph)
        g = make_gpu(case)
        snaps = []
        dt = case.dt
        for s in range(1, 6):
            g.navier_stokes_solver(s, dt, want_checks=False)
        snaps.append([g.download(f) for f in comps])
        g.run(5, 4, dt)
        snaps.append([g.download(f) for f in comps] + [g.scalars()["maxdiv"]])
        g.run(9, 3, 0.5 * dt)                      # new dt: captured again
        g.set_params(density=case.density, viscosity=2 * case.viscosity, g=case.g, cfl=case.cfl)
        g.run(12, 3, 0.5 * dt)                     # new parameters: captured again
        snaps.append([g.download(f) for f in comps] + [g.scalars()["maxdiv"]])
        results.append(snaps)
        g.close()
    for a, b in zip(*results):
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x), np.asarray(y))

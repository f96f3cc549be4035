"""The drop-in step (fen_ns_step, what FEN's navier_stokes_solver calls) with
its projection left pending and fused into the next step's momentum.

Every comparison is bitwise against a context created with FEN_FUSED_RUN=0,
which runs each step whole (momentum, solve, projection, checks): the
pending path must be invisible to every reader of the state and to every
call that changes it between steps.
"""
import numpy as np
import pytest

from fen_amd import Solver, cases
from fen_amd._lib import BC_WALL, DVO_X, DVO_Y, DVO_Z, FP32_REFINE, P, PHI, S_X, S_Y, S_Z, U, V, W
from tests.common import make_gpu

pytestmark = pytest.mark.gpu

FIELDS = (U, V, W, P, PHI, DVO_X, DVO_Y, DVO_Z)


def _pair(case, monkeypatch, precision=0):
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("FEN_FUSED_RUN", fused)
        out[fused] = make_gpu(case, precision=precision)
    monkeypatch.delenv("FEN_FUSED_RUN")
    return out["1"], out["0"]


def _same_state(a, b):
    for f in FIELDS:
        x, y = a.download(f), b.download(f)
        assert np.array_equal(x, y), (f, np.argwhere(x != y)[:4])
    sa, sb = a.scalars(), b.scalars()
    assert sa["maxdiv"] == sb["maxdiv"] and sa["maxcfl"] == sb["maxcfl"]


def _case(name, steps):
    if name == "tgv":
        return cases.tgv3d(32, steps=steps)
    if name == "channel":
        return cases.channel(32, 16, 24, steps=steps)
    if name == "cavity":   # nnn: walls on all six faces, moving lid
        return cases.cavity3d(32, 16, 24, steps=steps)
    if name == "duct":     # npn: walls at x and z
        return cases.duct3d(32, 16, 24, steps=steps)
    return cases.duct3d(32, 16, 24, steps=steps, z_io=True)   # npn with Inflow/Outflow at z


@pytest.mark.parametrize("precision,name", [(0, "tgv"), (0, "channel"), (0, "cavity"), (0, "duct"), (0, "duct_io"),
                                            (FP32_REFINE, "tgv"), (FP32_REFINE, "channel")])
def test_queued_steps_equal_whole_steps(name, precision, monkeypatch):
    print(f"dropin queued {name} precision {precision}", flush=True)
    case = _case(name, 6)
    a, b = _pair(case, monkeypatch, precision)
    for s in range(1, 7):
        a.navier_stokes_solver(s, case.dt, want_checks=False)
        b.navier_stokes_solver(s, case.dt, want_checks=False)
    _same_state(a, b)


@pytest.mark.parametrize("name", ["tgv", "channel", "cavity", "duct_io"])
def test_checks_read_every_step(name, monkeypatch):
    """maxdiv and maxCFL of every step from the pending state's checks equal
    those of the whole step."""
    case = _case(name, 5)
    a, b = _pair(case, monkeypatch)
    for s in range(1, 6):
        ra = a.navier_stokes_solver(s, case.dt)
        rb = b.navier_stokes_solver(s, case.dt)
        assert ra == rb, s
    _same_state(a, b)


def test_calls_between_steps(monkeypatch):
    """Reads and changes between pending steps: field downloads, a source
    upload, boundary planes, a ghost fill, a stand-alone Poisson solve, a
    parameter change; then more steps."""
    case = cases.channel(32, 16, 24, steps=12)
    a, b = _pair(case, monkeypatch)
    rng = np.random.default_rng(21)
    dt = case.dt
    src = {f: np.asfortranarray(0.01 * rng.standard_normal(case.n)) for f in (S_X, S_Y, S_Z)}
    plane = 0.02 * rng.standard_normal((34, 18))
    rhs = np.asfortranarray(rng.standard_normal(case.n))
    rhs -= rhs.mean()
    step = 0

    def both(fn):
        return fn(a), fn(b)

    def steps(k):
        nonlocal step
        for _ in range(k):
            step += 1
            both(lambda g: g.navier_stokes_solver(step, dt, want_checks=False))

    steps(2)
    x, y = both(lambda g: g.download(U))
    assert np.array_equal(x, y)
    steps(2)
    for f, arr in src.items():
        both(lambda g: g.upload(f, arr))
    steps(2)
    both(lambda g: g.set_bc_plane(U, 5, plane))
    steps(1)
    both(lambda g: g.update_ghost_nodes(P))
    steps(1)
    x, y = both(lambda g: g.solve_poisson(rhs, 0))
    assert np.array_equal(x, y)
    steps(1)
    both(lambda g: g.set_params(density=1.0, viscosity=case.viscosity * 1.5, g=case.g, cfl=case.cfl))
    steps(2)
    _same_state(a, b)


def test_constant_cfl_queued(monkeypatch):
    """Constant CFL: the dt of each step comes from the pending step's checks."""
    case = cases.tgv3d(32, steps=8)
    a, b = _pair(case, monkeypatch)
    for g in (a, b):
        g.set_params(density=case.density, viscosity=case.viscosity, cfl=0.7, constant_cfl=True)
        g.set_timestep(2.0)
    da = db = case.dt
    for s in range(1, 9):
        da, _, _ = a.navier_stokes_solver(s, da, want_checks=False)
        db, _, _ = b.navier_stokes_solver(s, db, want_checks=False)
        assert da == db, s
    _same_state(a, b)


def test_poisson_solve_keeps_phi(monkeypatch):
    """A stand-alone solve between steps leaves phi as the step left it
    (ppn: phi carries a pending mean shift)."""
    case = cases.channel(32, 16, 24, steps=3)
    g = make_gpu(case)
    assert g.solver == "ppn"
    for s in (1, 2, 3):
        g.navier_stokes_solver(s, case.dt, want_checks=False)
    phi0 = g.download(PHI)
    rng = np.random.default_rng(2)
    f = np.asfortranarray(rng.standard_normal(case.n))
    g.solve_poisson(f - f.mean(), 0)
    assert np.array_equal(g.download(PHI), phi0)


def test_run_then_steps(monkeypatch):
    """fen_ns_run followed by single steps, with a download in between."""
    case = cases.tgv3d(32, steps=9)
    a, b = _pair(case, monkeypatch)
    for g in (a, b):
        g.run(0, 4, case.dt)
    assert np.array_equal(a.download(P), b.download(P))
    for g in (a, b):
        g.run(4, 3, case.dt)
        for s in (8, 9):
            g.navier_stokes_solver(s, case.dt, want_checks=False)
    _same_state(a, b)

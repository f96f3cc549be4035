"""The restart file through the C-ABI (fen_save_state / fen_load_state,
reference save_state / load_state, src/solver.f90:160-329).

The file is the interiors of p, u, v, dv_o%x, dv_o%y [, w, dv_o%z]
[, vof, rho, mu, p_o] as raw float64 in Fortran order; it is checked
against the downloaded fields, and the reference's restart criterion
(test/small_test/io/run.sh:6-16: a run restarted from a saved step ends in
the same state file as the run that saved it) is checked bitwise for a
periodic, a channel, a cavity and a two-phase case.
"""
import math

import numpy as np
import pytest

from fen_amd import FenError, Solver, cases
from fen_amd._lib import DVO_X, DVO_Y, DVO_Z, MU, P, RHO, U, V, VOF, W
from tests.common import make_gpu

pytestmark = pytest.mark.gpu


def _interior(a, dim):
    return a[1:-1, 1:-1, 1:2] if dim == 2 else a[1:-1, 1:-1, 1:-1]


@pytest.mark.parametrize("name", ["tgv", "channel", "cavity"])
def test_state_file_layout_and_restart(name, tmp_path):
    case = {"tgv": lambda: cases.tgv3d(16, steps=10), "channel": lambda: cases.channel(32, 16, 24, steps=10),
            "cavity": lambda: cases.cavity3d(16, 16, 8, steps=10)}[name]()
    a = make_gpu(case)
    dt = case.dt
    for s in range(1, 6):
        a.navier_stokes_solver(s, dt, want_checks=False)
    a.save_state(str(tmp_path / "s5.raw"))
    # layout: the interiors in the reference's order
    n = case.n[0] * case.n[1] * case.n[2]
    raw = np.fromfile(tmp_path / "s5.raw")
    assert raw.size == 7 * n
    parts = [P, U, V, DVO_X, DVO_Y, W, DVO_Z]
    for q, f in enumerate(parts):
        d = a.download(f)
        want = _interior(d, 3) if f in (P, U, V, W) else d
        assert np.array_equal(raw[q * n:(q + 1) * n], want.ravel(order="F")), f
    for s in range(6, 11):
        a.navier_stokes_solver(s, dt, want_checks=False)
    a.save_state(str(tmp_path / "s10.raw"))
    b = make_gpu(case)
    b.load_state(str(tmp_path / "s5.raw"))
    for s in range(6, 11):
        b.navier_stokes_solver(s, dt, want_checks=False)
    b.save_state(str(tmp_path / "r10.raw"))
    assert (tmp_path / "s10.raw").read_bytes() == (tmp_path / "r10.raw").read_bytes()
    a.close()
    b.close()


def test_two_phase_state_and_restart(tmp_path):
    """Two-phase mode: vof, rho, mu and p_o follow the flow fields."""
    from tests.test_oracle_multiphase import vof_from_distance
    from fen_amd._lib import BC_PERIODIC, BC_WALL
    n = (16, 32)
    L = (1.0, 2.0)
    vof = vof_from_distance(lambda x, y: y - 1.0 - 0.05 * np.cos(2 * math.pi * x), n, L[0] / n[0])
    bc = [BC_PERIODIC, BC_PERIODIC, BC_WALL, BC_WALL, BC_PERIODIC, BC_PERIODIC]

    def make():
        g = Solver(2, (n[0], n[1], 1), (L[0], L[1], L[0] / n[0]), bc=bc)
        g.set_params(density=1.0, viscosity=0.01, g=(0.0, -1.0, 0.0), cfl=0.25)
        g.set_multiphase(1.0, 3.0, 0.01, 0.01, 0.0)
        full = np.zeros((n[0] + 2, n[1] + 2, 3), order="F")
        full[1:-1, 1:-1, 1] = vof
        g.upload(VOF, full)
        g.update_ghost_nodes(VOF)
        g.update_material_properties()
        return g, g.set_timestep(1.0)

    a, dt = make()
    for s in range(1, 6):
        a.navier_stokes_solver(s, dt)
    a.save_state(str(tmp_path / "s5.raw"))
    a5_x_first = a.x_first
    raw = np.fromfile(tmp_path / "s5.raw")
    m = n[0] * n[1]
    assert raw.size == 9 * m   # p, u, v, dv_o x2, vof, rho, mu, p_o
    for q, f in ((5, VOF), (6, RHO), (7, MU)):
        assert np.array_equal(raw[q * m:(q + 1) * m], a.download(f)[1:-1, 1:-1, 1].ravel(order="F")), f
    for s in range(6, 11):
        a.navier_stokes_solver(s, dt)
    a.save_state(str(tmp_path / "s10.raw"))
    b, _ = make()
    b.load_state(str(tmp_path / "s5.raw"))
    # the sweep order is not in the file: five advections flipped it, and the
    # restarted run sets it as the reference's restart test does
    # (test/small_test/io/test_MF.f90:75-80)
    assert b.x_first and not a5_x_first
    b.x_first = not b.x_first
    for s in range(6, 11):
        b.navier_stokes_solver(s, dt)
    b.save_state(str(tmp_path / "r10.raw"))
    assert (tmp_path / "s10.raw").read_bytes() == (tmp_path / "r10.raw").read_bytes()


def test_load_state_refuses_a_file_of_another_grid(tmp_path):
    a = make_gpu(cases.tgv3d(16, steps=1))
    a.save_state(str(tmp_path / "s.raw"))
    b = make_gpu(cases.tgv3d(32, steps=1))
    with pytest.raises(FenError, match="bytes; this grid's state is"):
        b.load_state(str(tmp_path / "s.raw"))
    with pytest.raises(FenError, match="cannot open"):
        b.load_state(str(tmp_path / "missing.raw"))

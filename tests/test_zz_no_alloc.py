"""fen_poisson_solve_noalloc: the allocation-free variant of
fen_poisson_solve (its block sums and mean live in a fixed device array).
Its results must equal fen_poisson_solve's bit for bit, on two calls in a
row with the second on a smaller grid (the array keeps the first call's
contents), for every solver kind and both ghost levels; it must leave a
pending step's phi alone and refuse what its array cannot hold."""
import numpy as np
import pytest

from fen_amd import FenError, Solver, cases
from fen_amd._lib import BC_PERIODIC, BC_WALL, FP32_REFINE, PHI
from tests.common import make_gpu

pytestmark = pytest.mark.gpu

P6, W = BC_PERIODIC, BC_WALL
KINDS = {   # dim, bc, large grid, small grid
    "ppp": (3, [P6] * 6, (64, 32, 32), (16, 8, 16)),
    "ppn": (3, [P6] * 4 + [W] * 2, (64, 64, 48), (16, 16, 8)),
    "npn": (3, [W, W, P6, P6, W, W], (32, 32, 24), (16, 8, 8)),
    "nnn": (3, [W] * 6, (32, 32, 16), (8, 16, 8)),
    "pp": (2, [P6] * 6, (128, 64, 1), (32, 16, 1)),
    "pn": (2, [P6, P6, W, W, P6, P6], (128, 64, 1), (16, 8, 1)),
    "nn": (2, [W, W, W, W, P6, P6], (64, 48, 1), (16, 8, 1)),
}


def _rhs(shape, gl, seed):
    rng = np.random.default_rng(seed)
    shp = tuple(s + 2 for s in shape) if gl else shape
    f = np.asfortranarray(rng.standard_normal(shp))
    return f - f.mean()


@pytest.mark.parametrize("kind", sorted(KINDS))
def test_noalloc_equals_solve_twice_second_smaller(kind):
    dim, bc, big, small = KINDS[kind]
    print(f"noalloc {kind} {big} then {small}", flush=True)
    for gl in (0, 1):
        for n, seed in ((big, 1), (small, 2)):
            s = Solver(dim, n, (1.0, n[1] / n[0], n[2] / n[0]), bc=bc)
            assert s.solver == kind
            f = _rhs(s.n, gl, seed + 10 * gl)
            want = s.solve_poisson(f, gl)
            got = s.solve_poisson_noalloc(f, gl)
            assert np.array_equal(got, want), (kind, n, gl, float(np.abs(got - want).max()))
            assert np.isfinite(got).all()
            s.close()


def test_noalloc_keeps_pending_phi():
    """ppn steps leave phi with its mean pending; the allocation-free solve
    uses its own mean slot, so phi reads back as the step left it."""
    case = cases.channel(32, 16, 24, steps=3)
    g = make_gpu(case)
    for s in (1, 2, 3):
        g.navier_stokes_solver(s, case.dt, want_checks=False)
    phi0 = g.download(PHI)
    for s in (4, 5):
        g.navier_stokes_solver(s, case.dt, want_checks=False)
    ref = make_gpu(case)
    for s in range(1, 6):
        ref.navier_stokes_solver(s, case.dt, want_checks=False)
    f = _rhs(g.n, 0, 7)
    a = g.solve_poisson_noalloc(f, 0)
    assert np.array_equal(a, ref.solve_poisson(f, 0))
    assert np.array_equal(g.download(PHI), ref.download(PHI))
    assert not np.array_equal(phi0, g.download(PHI))
    g.close()
    ref.close()


def test_noalloc_refuses_what_it_cannot_hold():
    s = Solver(3, (16, 16, 16), (1.0, 1.0, 1.0), precision=FP32_REFINE)
    with pytest.raises(FenError, match="FP64 contexts"):
        s.solve_poisson_noalloc(np.zeros((16, 16, 16), order="F"), 0)
    s.close()
    # ppn with nx = 16: 256 rows per c2r block, ny*nz/256 = 32800 block sums > 32768
    n = (16, 1024, 8200)
    s = Solver(3, n, (1.0, 64.0, 8200.0 / 16), bc=[P6] * 4 + [W] * 2)
    with pytest.raises(FenError, match="grid too large"):
        s.solve_poisson_noalloc(np.zeros(n, order="F"), 0)
    s.close()

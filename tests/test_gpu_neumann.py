"""The Poisson solvers with Neumann directions -- pn and nn in 2-D, npn and
nnn in 3-D (src/poisson.f90:306-596, 1177-1451) -- on the GPU against the
oracle, on identical seeded inputs.

Direct solves (fen_poisson_solve, both ghost levels) and short runs of the
time step on walled, driven and Inflow/Outflow cases.  The GPU's cosine
transforms go through complex FFTs (Makhoul's form), the oracle's follow
the definition, so results agree to rounding: the bar is 1e-12 relative to
the largest value of each field, as for the periodic solvers.
"""
import numpy as np
import pytest

from fen_amd import cases
from fen_amd._lib import BC_INFLOW, BC_OUTFLOW, BC_PERIODIC, BC_WALL, P, U, V, W
from oracle import oracle as O
from tests.common import live, make_gpu, make_oracle, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12
Pe, Wa, In, Out = BC_PERIODIC, BC_WALL, BC_INFLOW, BC_OUTFLOW

# kind, dim, n, bc; the Thomas direction has a length that is not a power of two where allowed
SOLVES = [
    ("pn", 2, (64, 32, 1), [Pe, Pe, Wa, Wa, Pe, Pe]),
    ("pn", 2, (32, 16, 1), [Pe, Pe, In, Out, Pe, Pe]),
    ("nn", 2, (32, 24, 1), [Wa, Wa, Wa, Wa, Pe, Pe]),
    ("nn", 2, (16, 40, 1), [Wa, Wa, In, Out, Pe, Pe]),
    ("npn", 3, (32, 16, 24), [Wa, Wa, Pe, Pe, Wa, Wa]),
    ("npn", 3, (16, 32, 20), [Wa, Wa, Pe, Pe, In, Out]),
    ("nnn", 3, (16, 16, 20), [Wa] * 6),
    ("nnn", 3, (32, 8, 16), [Wa, Wa, Wa, Wa, In, Out]),
]


@pytest.mark.parametrize("kind,dim,n,bc", SOLVES)
@pytest.mark.parametrize("gl", [0, 1])
def test_neumann_poisson_solve_matches_oracle(kind, dim, n, bc, gl):
    from fen_amd import Solver
    print(f"solve {kind} {n} bc {bc} gl {gl}", flush=True)
    rng = np.random.default_rng(17)
    L = (1.0, n[1] / n[0], n[2] / n[0])
    o = O.Oracle(dim, n, L, bc=bc)
    g = Solver(dim, n, L, bc=bc)
    assert o.solver == kind and g.solver == kind
    shp = (n[0] + 2 * gl, n[1] + 2 * gl, n[2] + 2 * gl)
    f = np.asfortranarray(rng.standard_normal(shp))
    fo = o.poisson_solve(f.copy(order="F"), gl)
    fg = g.solve_poisson(f, gl)
    e = rel_err(fg, fo)
    print(f"  rel err {e:.3e}")
    assert e < TOL, e
    g.close()
    o.close()


STEP_CASES = {
    "cavity2d_nn": lambda: cases.cavity2d(32, 24, steps=40),
    "channel2d_pn": lambda: cases.channel2d(64, 32, steps=40),
    "channel2d_pn_inflow": lambda: cases.channel2d(32, 32, inflow=True, steps=30),
    "duct3d_npn": lambda: cases.duct3d(32, 16, 24, steps=20),
    "duct3d_npn_inflow": lambda: cases.duct3d(16, 32, 20, z_io=True, steps=20),
    "cavity3d_nnn": lambda: cases.cavity3d(16, 16, 20, steps=20),
    # the DCT-II's register-loading form at the ends of its range of
    # transform lengths (x: 8, 128, 512; y: 8, 128) and with an odd row count
    "cavity3d_nnn_x8": lambda: cases.cavity3d(8, 8, 6, steps=10),
    "cavity3d_nnn_x128": lambda: cases.cavity3d(128, 16, 8, steps=6),
    "cavity3d_nnn_y128": lambda: cases.cavity3d(16, 128, 6, steps=6),
    "duct3d_npn_x512": lambda: cases.duct3d(512, 8, 6, steps=4),
    "cavity2d_nn_odd_ny": lambda: cases.cavity2d(64, 25, steps=20),
}


@pytest.mark.parametrize("name", list(STEP_CASES))
def test_neumann_steps_match_oracle(name):
    case = STEP_CASES[name]()
    print(f"steps {name} {case.n} x {case.steps}", flush=True)
    o = make_oracle(case)
    g = make_gpu(case)
    dt = case.dt
    for s in range(1, case.steps + 1):
        _, mdo, mco = o.step(s, dt)
        _, mdg, mcg = g.navier_stokes_solver(s, dt)
    assert abs(mcg - mco) <= 1e-12 * abs(mco)
    if name != "channel2d_pn_inflow":
        # pn removes the interior mean of phi (src/poisson.f90:408) also when
        # the Outflow face fixes phi = 0, so that case keeps a divergence
        assert abs(mdg) < 1e-12 and abs(mdo) < 1e-12, (mdg, mdo)
    else:
        assert abs(mdg - mdo) <= 1e-9 * abs(mdo), (mdg, mdo)
    comps = [U, V, P] if case.dim == 2 else [U, V, W, P]
    for fid in comps:
        e = rel_err(live(g.download(fid), case.dim), live(o.field(fid), case.dim))
        assert e < TOL, (fid, e)
    g.close()
    o.close()


@pytest.mark.parametrize("name", ["cavity2d_nn", "duct3d_npn_inflow"])
def test_neumann_step_graph_replay_matches_eager(name, monkeypatch):
    """Replays of the captured step equal eager steps bit for bit."""
    case = STEP_CASES[name]()
    comps = [U, V, P] if case.dim == 2 else [U, V, W, P]
    results = []
    for graph in ("0", "1"):
        monkeypatch.setenv("FEN_STEP_GRAPH", graph)
        g = make_gpu(case)
        g.run(0, 6, case.dt)
        results.append([g.download(f) for f in comps] + [g.scalars()["maxdiv"]])
        g.close()
    for x, y in zip(*results):
        assert np.array_equal(np.asarray(x), np.asarray(y))

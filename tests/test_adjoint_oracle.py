"""The oracle's discrete adjoint (the FP64 reference the GPU adjoint is
checked against) pinned by central finite differences of the oracle's own
forward step, on small triply periodic grids.

Loss: kinetic energy of the final velocity, 0.5 * sum(u^2 + v^2 + w^2) over
the interior.  The adjoint seeded with the final velocity returns the
gradient with respect to the initial velocity, pressure and dv_o; each is
compared with (L(x + eps d) - L(x - eps d)) / (2 eps) along random
directions d.
"""
import numpy as np
import pytest

from fen_amd._lib import DVO_X, DVO_Y, DVO_Z, P, U, V, W
from oracle import oracle as O

N = (16, 8, 8)
L = (2 * np.pi, np.pi, np.pi)


def _initial(seed=7):
    rng = np.random.default_rng(seed)
    nx, ny, nz = N
    x = (np.arange(1, nx + 1) - 0.5) * (L[0] / nx)
    y = (np.arange(1, ny + 1) - 0.5) * (L[1] / ny)
    z = (np.arange(1, nz + 1) - 0.5) * (L[2] / nz)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    u = np.sin(X) * np.cos(2 * Y) * np.cos(2 * Z) + 0.1 * rng.standard_normal(N)
    v = -np.cos(X) * np.sin(2 * Y) * np.cos(2 * Z) + 0.1 * rng.standard_normal(N)
    w = 0.1 * rng.standard_normal(N)
    p = 0.1 * rng.standard_normal(N)
    return {U: u, V: v, W: w, P: p, DVO_X: np.zeros(N), DVO_Y: np.zeros(N), DVO_Z: np.zeros(N)}


def _oracle(nu=1.0 / 400):
    o = O.Oracle(3, N, L)
    o.set_params(density=1.0, viscosity=nu, cfl=0.5)
    dt = o.set_timestep(2.0)
    return o, dt


def _load(o, state):
    for fid, arr in state.items():
        f = o.field(fid)
        if fid <= W:
            f[1:-1, 1:-1, 1:-1] = arr
            o.update_ghost_nodes(fid)
        else:
            f[...] = arr


def _forward(o, dt, state, nsteps, record=False):
    _load(o, state)
    o.set_dt_o(dt)
    states = []
    for s in range(1, nsteps + 1):
        if record:
            states.append([o.field(f).copy(order="F") for f in (U, V, W)])
        o.step(s, dt)
    final = [o.field(f)[1:-1, 1:-1, 1:-1].copy(order="F") for f in (U, V, W)]
    return final, states


def _loss_difference(fa, fb):
    """L(a) - L(b) as 0.5 * sum((a + b) * (a - b)): the difference of the two
    final states is formed first, so no digits cancel in the loss itself
    (forming L(a) - L(b) directly limits central differences to about
    eps_mach * |L| / eps)."""
    return 0.5 * sum(float(((a + b) * (a - b)).sum()) for a, b in zip(fa, fb))


def _gradient(o, dt, state, nsteps):
    final, states = _forward(o, dt, state, nsteps, record=True)
    ub, vb, wb = [np.asfortranarray(a.copy()) for a in final]
    pb = np.zeros(N, order="F")
    dvob = [np.zeros(N, order="F") for _ in range(3)]
    for s in reversed(range(nsteps)):
        for f, arr in zip((U, V, W), states[s]):
            o.field(f)[...] = arr
        o.adjoint_step(dt, ub, vb, wb, pb, dvob)
    return {U: ub, V: vb, W: wb, P: pb, DVO_X: dvob[0], DVO_Y: dvob[1], DVO_Z: dvob[2]}


@pytest.mark.parametrize("nsteps", [1, 6])
def test_oracle_adjoint_matches_central_differences(nsteps):
    o, dt = _oracle()
    state = _initial()
    grad = _gradient(o, dt, state, nsteps)
    rng = np.random.default_rng(11)
    eps = 1e-6
    for trial in range(4):
        # a direction over velocity, pressure and dv_o together
        d = {f: rng.standard_normal(N) for f in state}
        if nsteps > 1 and trial == 0:
            d = {f: (d[f] if f in (U, V, W) else np.zeros(N)) for f in d}   # velocity only
        plus = {f: state[f] + eps * d[f] for f in state}
        minus = {f: state[f] - eps * d[f] for f in state}
        fd = _loss_difference(_forward(o, dt, plus, nsteps)[0], _forward(o, dt, minus, nsteps)[0]) / (2 * eps)
        ad = sum(float((grad[f] * d[f]).sum()) for f in state)
        assert abs(ad - fd) <= 1e-6 * abs(fd), (trial, ad, fd)   # the bar fixed for the adjoint
    o.close()


def test_oracle_adjoint_rejects_non_periodic():
    o = O.Oracle(3, (8, 8, 8), (1.0, 1.0, 1.0), bc=[0, 0, 0, 0, 1, 1])
    a = [np.zeros((8, 8, 8), order="F") for _ in range(7)]
    with pytest.raises(ValueError):
        o.adjoint_step(1e-3, a[0], a[1], a[2], a[3], a[4:])
    o.close()

"""The GPU discrete adjoint (fen_adjoint_record / fen_adjoint_backward)
against the oracle's reference adjoint and against central differences of
the GPU forward run, through a torch.autograd.Function harness.

Loss: kinetic energy of the final velocity, 0.5 * sum(u^2 + v^2 + w^2).
Central differences use L(a) - L(b) = 0.5 * sum((a + b)(a - b)) over the two
final states, so no digits cancel in L itself.
"""
import numpy as np
import pytest
import torch

from fen_amd import cases
from fen_amd._lib import DVO_X, DVO_Y, DVO_Z, P, U, V, W
from tests import test_adjoint_oracle as ref
from tests.adjoint_harness import NSRun, Run, final_velocity, load_state

pytestmark = pytest.mark.gpu


def _solver(n, L, nu, cfl=0.5, Uref=2.0):
    from fen_amd import Solver
    s = Solver(3, n, L)
    s.set_params(density=1.0, viscosity=nu, cfl=cfl)
    dt = s.set_timestep(Uref)
    return s, dt


def _ke_grad(run, u0):
    x = [torch.tensor(a, requires_grad=True) for a in u0]
    out = NSRun.apply(x[0], x[1], x[2], run)
    loss = 0.5 * sum((o * o).sum() for o in out)
    loss.backward()
    return [t.grad.numpy() for t in x]


@pytest.mark.parametrize("shape", [None, (16, 8, 32)])
def test_gpu_adjoint_matches_oracle_adjoint(shape, monkeypatch):
    # (16, 8, 32): the transposed momentum marches z in two chunks of planes
    if shape is not None:
        monkeypatch.setattr(ref, "N", shape)
        monkeypatch.setattr(ref, "L", (2 * np.pi, np.pi, 4 * np.pi))
    o, dt = ref._oracle()
    state = ref._initial()
    nsteps = 6
    g_ref = ref._gradient(o, dt, state, nsteps)
    s, dt_g = _solver(ref.N, ref.L, 1.0 / 400)
    assert dt_g == dt
    load_state(s, state, dt)
    s.adjoint_record(0, nsteps, 2, dt)
    seed = [np.asfortranarray(a) for a in final_velocity(s)]
    g = s.adjoint_backward(seed, want_p=True)
    for got, fid in zip(g[:3], (U, V, W)):
        err = np.abs(got - g_ref[fid]).max() / np.abs(g_ref[fid]).max()
        print(f"field {fid}: GPU adjoint vs oracle adjoint, max rel {err:.2e}")
        assert err <= 1e-10
    # p enters u* only as a gradient, which the projection removes exactly
    # (D G S = I off the zero mode), so the final velocity does not depend on
    # p0: dL/dp0 vanishes and both adjoints hold round-off only.
    scale = max(np.abs(a).max() for a in g[:3])
    print(f"dL/dp0: GPU max {np.abs(g[3]).max():.2e}, oracle max {np.abs(g_ref[P]).max():.2e} (velocity scale {scale:.2e})")
    assert np.abs(g[3]).max() <= 1e-12 * scale and np.abs(g_ref[P]).max() <= 1e-12 * scale
    o.close()
    s.close()


def test_gpu_adjoint_checkpoint_interval_is_exact():
    """Recomputation from checkpoints replays the forward bit for bit."""
    state = ref._initial(seed=3)
    s, dt = _solver(ref.N, ref.L, 1.0 / 400)
    grads = []
    for every in (1, 3, 7):
        load_state(s, state, dt)
        s.adjoint_record(0, 7, every, dt)
        seed = [np.asfortranarray(a) for a in final_velocity(s)]
        grads.append(s.adjoint_backward(seed))
    for g in grads[1:]:
        for a, b in zip(g, grads[0]):
            assert np.array_equal(a, b)
    s.close()


def test_gpu_adjoint_fused_recompute_is_exact(monkeypatch):
    """The backward sweep recomputes each segment as a fused run whose
    kernel saves every state it forms; with FEN_FUSED_RUN=0 it steps one
    step at a time into the segment buffer instead.  Both give the same
    gradient bit for bit."""
    state = ref._initial(seed=5)
    grads = []
    for fused in ("1", "0"):
        monkeypatch.setenv("FEN_FUSED_RUN", fused)
        s, dt = _solver(ref.N, ref.L, 1.0 / 400)
        load_state(s, state, dt)
        s.adjoint_record(0, 9, 4, dt)
        seed = [np.asfortranarray(a) for a in final_velocity(s)]
        grads.append(s.adjoint_backward(seed))
        s.close()
    for a, b in zip(grads[0], grads[1]):
        assert np.array_equal(a, b)


def test_gpu_adjoint_tgv128_re400_matches_central_differences():
    """Discrete adjoint of the 3-D periodic 128^3 case, Re = 400, 200 steps,
    checkpoints every 20: d(KE_final)/d(u0) along 16 random directions
    against central differences with eps = 1e-6, to 1e-6 relative."""
    case = cases.tgv3d(128, Re=400.0, steps=200)
    s, dt = _solver(case.n, case.L, case.viscosity)
    assert abs(dt - case.dt) <= 1e-15 * case.dt
    state = dict(case.fields)
    run = Run(s, state, 200, 20, dt, dt)
    u0 = [np.ascontiguousarray(state[f]) for f in (U, V, W)]
    grad = _ke_grad(run, u0)
    rng = np.random.default_rng(2026)
    eps = 1e-6
    worst = 0.0
    for trial in range(16):
        d = [rng.standard_normal(case.n) for _ in range(3)]
        finals = []
        for sgn in (1.0, -1.0):
            st = dict(state)
            for f, a, da in zip((U, V, W), u0, d):
                st[f] = a + sgn * eps * da
            load_state(s, st, dt)
            s.run(0, 200, dt)
            finals.append(final_velocity(s))
        fd = 0.5 * sum(float(((a + b) * (a - b)).sum()) for a, b in zip(*finals)) / (2 * eps)
        ad = sum(float((g * da).sum()) for g, da in zip(grad, d))
        rel = abs(ad - fd) / abs(fd)
        worst = max(worst, rel)
        assert rel <= 1e-6, (trial, ad, fd)
    print(f"tgv3d 128^3 Re 400, 200 steps, checkpoints every 20: adjoint vs central differences "
          f"(16 directions, eps 1e-6): worst relative difference {worst:.2e}")
    s.close()

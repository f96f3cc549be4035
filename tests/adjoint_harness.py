"""torch.autograd wrapper of a recorded run -- a test harness only.

NSRun.apply(u0, v0, w0, run) advances `run.nsteps` time steps of a libfen
Solver from (u0, v0, w0) and the rest of `run.state`, recording the run
through the C-ABI (fen_adjoint_record); its backward calls
fen_adjoint_backward with the incoming gradient as the seed.  Tensors are
CPU float64 with FEN's (i, j, k) indexing; the computation itself runs in
libfen on the GPU.
"""
import numpy as np
import torch

from fen_amd._lib import U, V, W
from fen_amd.cases import interior_velocity as final_velocity
from fen_amd.cases import load_state


class Run:
    def __init__(self, solver, state, nsteps, every, dt, dt_o):
        self.solver, self.state, self.nsteps, self.every, self.dt, self.dt_o = solver, state, nsteps, every, dt, dt_o


class NSRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, v0, w0, run):
        st = dict(run.state)
        st[U], st[V], st[W] = (t.detach().numpy() for t in (u0, v0, w0))
        load_state(run.solver, st, run.dt_o)
        run.solver.adjoint_record(0, run.nsteps, run.every, run.dt)
        ctx.run = run
        return tuple(torch.from_numpy(a) for a in final_velocity(run.solver))

    @staticmethod
    def backward(ctx, gu, gv, gw):
        seed = [np.asfortranarray(g.detach().numpy()) for g in (gu, gv, gw)]
        grads = ctx.run.solver.adjoint_backward(seed)
        return tuple(torch.from_numpy(np.ascontiguousarray(g)) for g in grads) + (None,)

"""Failure behaviour of the step on the GPU.

The reference has no check for a non-finite field; the C-ABI reports one
through the step's own checks (the max-divergence / max-velocity reduction
of navier_stokes_solver, src/navier_stokes.f90:570-619) with
FEN_ERR_NONFINITE, on the single-step path and after a fused run.  A NaN is
ordinary data for every kernel: nothing faults, the run just carries it.
"""
import math

import numpy as np
import pytest

from fen_amd import FenError, Solver, cases
from fen_amd._lib import P, U, V, W
from tests.common import make_gpu

pytestmark = pytest.mark.gpu

NONFINITE = -4


def _poisoned_case(N=32):
    case = cases.tgv3d(N, steps=4)
    u = case.fields[U].copy(order="F")
    u[N // 2, 3, 5] = math.nan
    case.fields[U] = u
    return case


def test_step_reports_nonfinite():
    g = make_gpu(_poisoned_case())
    with pytest.raises(FenError) as e:
        g.navier_stokes_solver(1, 1e-3)
    assert e.value.code == NONFINITE
    # the scalars stay non-finite on a later read
    with pytest.raises(FenError) as e:
        g.scalars()
    assert e.value.code == NONFINITE


def test_queued_steps_report_nonfinite_at_the_read():
    g = make_gpu(_poisoned_case())
    for s in (1, 2, 3):
        g.navier_stokes_solver(s, 1e-3, want_checks=False)
    with pytest.raises(FenError) as e:
        g.scalars()
    assert e.value.code == NONFINITE


def test_fused_run_reports_nonfinite():
    g = make_gpu(_poisoned_case())
    g.run(0, 4, 1e-3)        # 4 steps: the fused projection + momentum path
    with pytest.raises(FenError) as e:
        g.scalars()
    assert e.value.code == NONFINITE


def test_inf_in_pressure_reports_nonfinite():
    case = cases.tgv3d(32, steps=2)
    p = case.fields[P].copy(order="F")
    p[1, 2, 3] = math.inf
    case.fields[P] = p
    g = make_gpu(case)
    with pytest.raises(FenError) as e:
        g.navier_stokes_solver(1, 1e-3)
    assert e.value.code == NONFINITE


def test_clean_state_after_reupload():
    """A fresh upload of finite fields clears the condition."""
    case = _poisoned_case()
    g = make_gpu(case)
    with pytest.raises(FenError):
        g.navier_stokes_solver(1, 1e-3)
    clean = cases.tgv3d(32, steps=1)
    for fid in (U, V, W, P):
        from tests.common import ghosted
        g.upload(fid, ghosted(clean.fields[fid], 3))
        g.update_ghost_nodes(fid)
    g.upload(5, np.zeros((32, 32, 32), order="F"))
    g.upload(6, np.zeros((32, 32, 32), order="F"))
    g.upload(7, np.zeros((32, 32, 32), order="F"))
    dt, md, mc = g.navier_stokes_solver(1, 1e-3)
    assert abs(md) < 1e-12 and math.isfinite(mc)

"""Constant-CFL stepping on the GPU against the oracle.

update_timestep (reference src/navier_stokes.f90:670-730) runs at the top of
every step when constant_CFL is set (:78): the max of |u|+|v|+|w| over the
interior sets dt, capped at 1.1 dt_o, and AB2 uses the variable-dt
coefficients.  Two cases:

* the uniform-field scenarios whose closed form pins the oracle
  (tests/test_timestep.py): no transform rounding enters, so the dt sequence
  and the fields must be bitwise equal;
* the reference's forced isotropic turbulence setup
  (test/large_test/isotropic_turbulence/isotropic.f90:30-58,115-150):
  128^3, CFL 0.9, the forcing S = 0.1 (u - mean u) recomputed from the
  velocity before every step and uploaded, 20 steps.  Velocities go through
  the FFT solve, which rounds differently on the two sides, so dt agrees to
  round-off (1e-13 relative per step) and the fields to 1e-12.
"""
import numpy as np
import pytest

from fen_amd import cases
from fen_amd._lib import P, S_X, S_Y, S_Z, U, V, W
from oracle import oracle as O
from tests.common import make_gpu, make_oracle, rel_err
from tests.test_timestep import closed_form_run, uniform_cfl_scenarios

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sc", uniform_cfl_scenarios(), ids=lambda s: s["name"])
def test_constant_cfl_uniform_bitwise(sc):
    from fen_amd import Solver
    rho, cfl = 2.0, 0.5
    n, L = (8, 8, 8), (1.0, 1.0, 1.0)
    o = O.Oracle(3, n, L)
    g = Solver(3, n, L)
    for s in (o, g):
        s.set_params(density=rho, viscosity=sc["nu"], cfl=cfl)
    for fid, val in zip((U, V, W), sc["u0"]):
        o.field(fid)[...] = val
        g.upload(fid, np.full((10, 10, 10), val, order="F"))
    dt_o_, dt_g = o.set_timestep(sc["U"]), g.set_timestep(sc["U"])
    assert dt_o_ == dt_g
    for s in (o, g):
        s.set_params(density=rho, viscosity=sc["nu"], cfl=cfl, constant_cfl=True)
    ref = closed_form_run(sc["u0"], sc["S"], rho, cfl, 0.125, o.dt_limits()[0], dt_o_)
    dto, dtg = dt_o_, dt_g
    for step, s in enumerate(sc["S"], start=1):
        for fid, val in zip((S_X, S_Y, S_Z), s):
            o.field(fid)[...] = val
            g.upload(fid, np.full(n, val, order="F"))
        dto, mdo, mco = o.step(step, dto)
        dtg, mdg, mcg = g.navier_stokes_solver(step, dtg)
        assert dtg == dto == ref[step - 1][0], step
        assert g.scalars()["dt_o"] == o.dt_o
        assert g.scalars()["dt_conv"] == o.dt_limits()[1]
        assert mcg == mco and mdg == mdo
    for fid in (U, V, W, P):
        assert np.array_equal(g.download(fid)[1:-1, 1:-1, 1:-1], o.field(fid)[1:-1, 1:-1, 1:-1]), fid


@pytest.mark.timeout(600)
def test_isotropic_128_constant_cfl_matches_oracle():
    case = cases.isotropic(128)
    o = make_oracle(case)
    g = make_gpu(case)
    dts = []
    for s in (o, g):
        dts.append(s.set_timestep(case.set_timestep_U))
        s.set_params(density=case.density, viscosity=case.viscosity, cfl=case.cfl_run, constant_cfl=True)
    assert dts[0] == dts[1]
    dto = dtg = dts[0]
    seq_o, seq_g = [], []
    for step in range(1, 21):
        fo = cases.isotropic_forcing(*(o.field(f)[1:-1, 1:-1, 1:-1] for f in (U, V, W)))
        fg = cases.isotropic_forcing(*(g.download(f)[1:-1, 1:-1, 1:-1] for f in (U, V, W)))
        for fid, a, b in zip((S_X, S_Y, S_Z), fo, fg):
            o.field(fid)[...] = a
            g.upload(fid, np.asfortranarray(b))
        dto, mdo, mco = o.step(step, dto)
        dtg, mdg, mcg = g.navier_stokes_solver(step, dtg)
        seq_o.append(dto)
        seq_g.append(dtg)
        assert abs(dtg - dto) <= 1e-13 * dto, (step, dtg, dto)
        assert abs(mdg) < 1e-12 and abs(mdo) < 1e-12
        assert mcg == pytest.approx(mco, rel=1e-12)
    # the run is CFL-limited: dt changed from step to step
    assert len(set(seq_o)) > 10
    for fid in (U, V, W, P):
        e = rel_err(g.download(fid)[1:-1, 1:-1, 1:-1], o.field(fid)[1:-1, 1:-1, 1:-1])
        assert e <= 1e-12, (fid, e)

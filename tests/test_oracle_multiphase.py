"""The oracle's two-phase (MF build) solver, pinned by what the mathematics
fixes (the reference's own VOF and multiphase tests are plots):

* MTHINC advection (src/volume_of_fluid.f90:420-544) in a divergence-free
  periodic flow conserves the phase integrals to round-off and stays in
  [0, 1]; a single vortex reversed at half time returns the circle;
  a planar interface translated by whole cells reproduces the shifted field.
* The reconstruction (:223-383): on a circle of radius R (vof = 1 outside)
  the normal is the outward radial direction and curv = -div n = -1/R.
* The MF step with equal phases (rho0 = rho1 = 1, mu0 = mu1, sigma = 0)
  reduces to the single-phase step: the Dodd & Ferrante terms cancel and the
  variable-viscosity stress div(2 mu D) equals mu lap(u) on a
  divergence-free field.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

RP = 0.5 * (1.0 + 1.0 / math.sqrt(3.0))
RM = 0.5 * (1.0 - 1.0 / math.sqrt(3.0))


def vof_from_distance(dist, n, delta, beta=1.0):
    """get_vof_from_distance (src/volume_of_fluid.f90:665-706) on cell centres."""
    i = np.arange(1, n[0] + 1)
    j = np.arange(1, n[1] + 1)
    X, Y = np.meshgrid((i - 0.5) * delta, (j - 0.5) * delta, indexing="ij")
    xm, xp = X + delta * (RM - 0.5), X + delta * (RP - 0.5)
    ym, yp = Y + delta * (RM - 0.5), Y + delta * (RP - 0.5)
    t = lambda x, y: 0.5 * (1.0 + np.tanh(beta * dist(x, y) / delta))
    return 0.5 * (0.5 * (t(xm, ym) + t(xp, ym)) + 0.5 * (t(xm, yp) + t(xp, yp)))


def _vof_oracle(n, L, bc=None):
    o = O.Oracle(2, (n[0], n[1], 1), (L[0], L[1], L[0] / n[0]), bc=bc)
    o.set_multiphase(1.0, 1.0, 1.0, 1.0)
    return o


def _set(o, fid, interior):
    o.field(fid)[1:-1, 1:-1, 1] = interior
    o.update_ghost_nodes(fid)


def _vortex(o, n, sign=1.0):
    d = o.delta
    i = np.arange(1, n + 1)
    _set(o, O.U, sign * np.sin(i[:, None] * d) * np.cos((i[None, :] - 0.5) * d))
    _set(o, O.V, -sign * np.cos((i[:, None] - 0.5) * d) * np.sin(i[None, :] * d))


def test_vof_advection_conserves_phase_and_stays_bounded():
    n = 64
    o = _vof_oracle((n, n), (2 * math.pi, 2 * math.pi))
    circle = lambda x, y: np.sqrt((x - math.pi) ** 2 + (y - 0.8 * math.pi) ** 2) - 0.4 * math.pi
    vof0 = vof_from_distance(circle, (n, n), o.delta)
    _set(o, O.VOF, vof0)
    _vortex(o, n)
    for _ in range(100):
        o.advect_vof(0.02)
    f = o.field(O.VOF)[1:-1, 1:-1, 1]
    assert abs(f.sum() - vof0.sum()) < 1e-13 * vof0.sum()
    assert abs((1 - f).sum() - (1 - vof0).sum()) < 1e-12 * (1 - vof0).sum()
    assert f.min() > -1e-12 and f.max() < 1 + 1e-12


def test_reversed_vortex_returns_the_circle():
    """test/small_test/volume_of_fluid/reversed: a circle stretched by a
    vortex and brought back by the reversed field (64^2, 2 x 160 steps)."""
    n = 64
    o = _vof_oracle((n, n), (2 * math.pi, 2 * math.pi))
    circle = lambda x, y: np.sqrt((x - math.pi) ** 2 + (y - 0.7 * math.pi) ** 2) - 0.4 * math.pi
    vof0 = vof_from_distance(circle, (n, n), o.delta)
    _set(o, O.VOF, vof0)
    _vortex(o, n)
    for _ in range(160):
        o.advect_vof(0.01)
    mid = o.field(O.VOF)[1:-1, 1:-1, 1].copy()
    _vortex(o, n, -1.0)
    for _ in range(160):
        o.advect_vof(0.01)
    f = o.field(O.VOF)[1:-1, 1:-1, 1]
    err = np.abs(f - vof0).sum() / np.abs(vof0 - 1).sum()
    moved = np.abs(mid - vof0).sum() / np.abs(vof0 - 1).sum()
    print(f"reversed vortex: relative L1 error {err:.3e} (after the forward half: {moved:.3e})")
    assert moved > 0.1          # the interface really moved
    assert err < 0.02 * moved


def test_planar_interface_translated_by_whole_cells():
    """A slab |x - 1/2| < 0.2 (two planar interfaces, periodic in x) moved
    4 cells by a uniform u reproduces the shifted field."""
    n = 32
    o = _vof_oracle((n, n), (1.0, 1.0))
    d = o.delta
    slab = lambda x, y: np.abs(x - 0.5) - 0.2 + 0.0 * y
    vof0 = vof_from_distance(slab, (n, n), d)
    _set(o, O.VOF, vof0)
    _set(o, O.U, np.full((n, n), 1.0))
    _set(o, O.V, np.zeros((n, n)))
    dt = 0.25 * d
    for _ in range(16):          # 4 cells to the right
        o.advect_vof(dt)
    f = o.field(O.VOF)[1:-1, 1:-1, 1]
    shifted = np.roll(vof0, 4, axis=0)
    print(f"translated slab: max difference {np.abs(f - shifted).max():.3e}")
    assert np.abs(f - shifted).max() < 2e-2
    assert abs(f.sum() - vof0.sum()) < 1e-12 * vof0.sum()


def test_reconstruction_normals_and_curvature():
    """A circle of radius R (vof = 1 outside): in the interface band the
    normal is the outward radial direction and the curvature is -1/R."""
    n = 64
    o = _vof_oracle((n, n), (1.0, 1.0))
    d = o.delta
    R = 0.25
    circle = lambda x, y: np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2) - R
    _set(o, O.VOF, vof_from_distance(circle, (n, n), d))
    o.vof_reconstruct()
    f = o.field(O.VOF)[1:-1, 1:-1, 1]
    band = (f > 0.2) & (f < 0.8)
    i = np.arange(1, n + 1)
    X, Y = np.meshgrid((i - 0.5) * d, (i - 0.5) * d, indexing="ij")
    ang = np.arctan2(Y - 0.5, X - 0.5)
    nx, ny = o.vof_aux(3)[1:-1, 1:-1, 1], o.vof_aux(4)[1:-1, 1:-1, 1]
    err = np.abs(np.angle(np.exp(1j * (np.arctan2(ny, nx) - ang))))
    curv = o.vof_aux(2)[1:-1, 1:-1, 1]
    mean = float(curv[band].mean())
    print(f"circle R = {R}: normal angle error max {err[band].max():.3e} rad, mean curvature {mean:.4f}")
    assert err[band].max() < 0.05
    assert abs(mean - (-1.0 / R)) < 0.05 / R


def test_mf_step_with_equal_phases_is_the_single_phase_step():
    n = 32
    L = 2 * math.pi
    nu = 0.05

    def make(mf):
        o = O.Oracle(2, (n, n, 1), (L, L, L / n))
        o.set_params(density=1.0, viscosity=nu, cfl=0.5)
        if mf:
            o.set_multiphase(1.0, 1.0, nu, nu, 0.0)
            _set(o, O.VOF, np.ones((n, n)))   # one phase: an intermediate uniform vof has no interface to reconstruct
            o.update_material_properties()
        d = o.delta
        i = np.arange(1, n + 1)
        _set(o, O.U, np.sin(i[:, None] * d) * np.cos((i[None, :] - 0.5) * d))
        _set(o, O.V, -np.cos((i[:, None] - 0.5) * d) * np.sin(i[None, :] * d))
        return o

    a, b = make(False), make(True)
    dt = 0.01
    a.set_dt_o(dt)
    b.set_dt_o(dt)
    for s in range(1, 21):
        a.step(s, dt)
        b.step(s, dt)
    for fid in (O.U, O.V, O.P):
        x, y = a.field(fid)[1:-1, 1:-1, 1], b.field(fid)[1:-1, 1:-1, 1]
        assert np.abs(x - y).max() < 1e-10 * np.abs(x).max(), fid


def test_mf_rejects_3d():
    o = O.Oracle(3, (8, 8, 8), (1.0, 1.0, 1.0))
    with pytest.raises(ValueError):
        o.set_multiphase(1.0, 2.0, 1.0, 1.0)


@pytest.mark.parametrize("rho_out", [1.0, 10.0])
def test_static_drop_laplace_pressure_jump(rho_out):
    """A circular drop at rest with surface tension sigma (CSF): the pressure
    inside exceeds the outside by sigma / R (2-D Laplace law), the spurious
    currents stay small and the velocity divergence-free; with a density
    ratio of 10 the Dodd & Ferrante splitting carries the variable density."""
    n = 64
    R, sigma = 0.25, 1.0
    o = O.Oracle(2, (n, n, 1), (1.0, 1.0, 1.0 / n))
    o.set_params(density=1.0, viscosity=0.1, cfl=0.5)
    o.set_multiphase(1.0, rho_out, 0.1, 0.1, sigma)       # phase 0 inside (vof = 0), phase 1 outside
    circle = lambda x, y: np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2) - R
    _set(o, O.VOF, vof_from_distance(circle, (n, n), o.delta))
    o.update_material_properties()
    dt = o.mf_set_timestep(1.0)
    for s in range(1, 201):
        _, md, _ = o.step(s, dt)
    p = o.field(O.P)[1:-1, 1:-1, 1]
    f = o.field(O.VOF)[1:-1, 1:-1, 1]
    jump = float(p[f < 0.01].mean() - p[f > 0.99].mean())
    umax = float(np.abs(o.field(O.U)[1:-1, 1:-1, 1]).max())
    print(f"static drop rho_out {rho_out}: jump {jump:.4f} (sigma/R = {sigma / R}), max|u| {umax:.2e}, maxdiv {md:.1e}")
    assert abs(jump - sigma / R) < 0.03 * sigma / R
    assert umax < 1e-2
    assert abs(md) < 1e-12


def _rt_oracle(n, L, sigma, cfl, constant_cfl):
    from fen_amd._lib import BC_PERIODIC, BC_WALL, VOF
    wave = lambda x, y: (y - 0.5 * L[1] - 0.05 * np.cos(2 * math.pi * x))
    vof = vof_from_distance(wave, n, L[0] / n[0])
    bc = [BC_PERIODIC, BC_PERIODIC, BC_WALL, BC_WALL, BC_PERIODIC, BC_PERIODIC]
    o = O.Oracle(2, (n[0], n[1], 1), (L[0], L[1], L[0] / n[0]), bc=bc)
    o.set_params(density=1.0, viscosity=0.01, g=(0.0, -1.0, 0.0), cfl=cfl)
    o.set_multiphase(1.0, 3.0, 0.01, 0.01, sigma)
    o.field(VOF)[1:-1, 1:-1, 1] = vof
    o.update_ghost_nodes(VOF)
    o.update_material_properties()
    dt = o.mf_set_timestep(1.0)
    o.set_params(density=1.0, viscosity=0.01, g=(0.0, -1.0, 0.0), cfl=cfl, constant_cfl=constant_cfl)
    return o, dt


def test_mf_constant_cfl_dt_sequence():
    """update_timestep's MF branch (src/navier_stokes.f90:670-730): before
    every step dt_o <- dt, dt = min(CFL delta / max(|u|+|v|), dt_visc,
    dt_surf), capped at 1.1 dt_o; the max is recomputed here from the
    oracle's velocity, dt_visc and dt_surf from set_timestep's closed forms
    (:653-660).  A small CFL makes dt_conv bind once the flow starts, the
    1.1 dt_o cap binds first (set_timestep's dt is for U = 1, the shear flow
    given here moves at 0.5)."""
    from fen_amd._lib import U, V
    n, L, sigma, cfl = (16, 32), (1.0, 2.0), 0.05, 0.02
    o, dt = _rt_oracle(n, L, sigma, cfl, True)
    d = L[0] / n[0]
    y = (np.arange(1, n[1] + 1) - 0.5) * d
    o.field(U)[1:-1, 1:-1, 1] = 0.5 * np.sin(math.pi * y / L[1])[None, :]   # u(y), v = 0: divergence-free
    o.update_ghost_nodes(U)
    dt_visc = 0.125 * d * d * min(1.0 / 0.01, 3.0 / 0.01)
    dt_surf = math.sqrt(0.5 * (1.0 + 3.0) * d ** 3 / (math.pi * sigma + 1.0e-16))
    assert dt == min(cfl * d / 1.0, dt_visc, dt_surf)
    bound = set()
    for s in range(1, 41):
        u, v = o.field(U)[1:-1, 1:-1, 1], o.field(V)[1:-1, 1:-1, 1]
        mv = float((np.abs(u) + np.abs(v)).max())
        dt_conv = cfl * d / mv if mv > 0 else 1.0
        want = min(dt_conv, dt_visc, dt_surf)
        cap = 1.1 * dt
        if want > cap:
            want = cap
        bound.add("cap" if want == cap else ("conv" if want == dt_conv else ("surf" if want == dt_surf else "visc")))
        dt_new, _, _ = o.step(s, dt)
        assert dt_new == want, (s, dt_new, want)
        assert o.dt_limits()[1] == dt_conv
        dt = dt_new
    assert {"cap", "conv"} <= bound, bound
    assert dt < 0.5 * dt_surf


def test_mf_constant_cfl_with_dt_fixed_by_dt_surf_is_the_fixed_dt_run():
    """When dt_surf binds on every step, dt stays equal to dt_o and the
    constant-CFL extrapolation p_o + (dt + dt_o)(p - p_o)/dt_o (:89) is
    2 p - p_o (:91) up to rounding: the run equals the fixed-dt run at that
    dt to round-off (a wrong dt or dt_o in the formula would not)."""
    from fen_amd._lib import P, U, V, VOF
    n, L = (16, 32), (1.0, 2.0)
    a, dta = _rt_oracle(n, L, 50.0, 1.0, True)
    b, dtb = _rt_oracle(n, L, 50.0, 1.0, False)
    assert dta == dtb
    for s in range(1, 11):
        dta, _, _ = a.step(s, dta)
        b.step(s, dtb)
        assert dta == dtb
        if s in (2, 10):   # one rounding of p_hat apart after step 2, grown by the flow after 10
            tol = 1e-14 if s == 2 else 1e-11
            for f in (U, V, P, VOF):
                x, y = a.field(f)[1:-1, 1:-1, 1], b.field(f)[1:-1, 1:-1, 1]
                assert np.abs(x - y).max() <= tol * np.abs(y).max(), (s, f)

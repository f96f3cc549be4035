"""FEN's Fortran interface over libfen: the layer compiles for DIM=2 and
DIM=3 and in the two-phase build (CPU), and FEN-style programs built on it
reproduce the oracle and the reference's restart criterion on the GPU."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from fen_amd import cases
from fen_amd._lib import P, U, V, W
from tests.common import live, make_oracle, rel_err

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "fen_amd", "fortran")
FC = "/opt/rocm/bin/amdflang"


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not installed")
@pytest.mark.parametrize("dim,mf", [(2, False), (3, False), (2, True)])
def test_fortran_layer_compiles(dim, mf, tmp_path):
    flags = ["-cpp", f"-DDIM={dim}"] + (["-DMF"] if mf else [])
    for src in ("fen_capi.f90", "fen_modules.F90"):
        subprocess.run([FC] + flags + ["-c", os.path.join(FDIR, src)], cwd=tmp_path, check=True)
    assert (tmp_path / "navier_stokes_mod.mod").exists() and (tmp_path / "poisson_mod.mod").exists()
    assert (tmp_path / "volume_of_fluid_mod.mod").exists() == mf
    assert (tmp_path / "multiphase_mod.mod").exists() == mf


def _bin(name):
    path = os.path.join(FDIR, "bin", name)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run __graft_entry__.build())")
    return path


@pytest.mark.gpu
def test_fortran_taylor_green_2d_matches_oracle(tmp_path):
    N, steps = 128, 1000
    (tmp_path / "data").mkdir()
    out = subprocess.run([_bin("taylor_green_2d"), str(N), str(steps)], cwd=tmp_path, check=True,
                         capture_output=True, text=True, timeout=120).stdout
    # print_navier_stokes_solver_status format (src/navier_stokes.f90:746)
    line = [l for l in out.splitlines() if l.startswith("step:")][-1]
    # (A6,I7,1x,A6,E13.6,1x,A4,E13.6,1x,A8,E13.6,1x,A9,1x,E13.6): 97 columns
    assert len(line) == 97, line
    assert line[:6] == "step: " and int(line[6:13]) == steps and line[14:20] == "time: "
    assert line[34:38] == "dt: " and line[52:60] == "maxdiv: " and line[74:83] == " maxCFL: "
    for a, b in ((20, 33), (38, 51), (60, 73), (84, 97)):
        assert re.match(r"^ *-?0\.\d{6}E[+-]\d\d$", line[a:b]), (line, line[a:b])
    case = cases.tgv2d(N, steps=steps)
    o = make_oracle(case)
    for s in range(1, steps + 1):
        o.step(s, case.dt)
    for name, fid in (("u", U), ("v", V), ("p", P)):
        got = np.fromfile(tmp_path / f"{name}.raw").reshape((N, N), order="F")
        ref = o.field(fid)[1:-1, 1:-1, 1]
        assert rel_err(got, ref) < 1e-12, name
    # save_fields: face values averaged to centres, 0.5*(f(i) + f(i-1)) (src/fields.f90:210-250)
    u, v = o.field(U)[:, :, 1], o.field(V)[:, :, 1]
    centred = {"vx": 0.5 * (u[1:-1, 1:-1] + u[:-2, 1:-1]), "vy": 0.5 * (v[1:-1, 1:-1] + v[1:-1, :-2]),
               "p": o.field(P)[1:-1, 1:-1, 1]}
    for name, ref in centred.items():
        got = np.fromfile(tmp_path / "data" / f"{name}_{steps:07d}.raw").reshape((N, N), order="F")
        assert rel_err(got, ref) < 1e-12, name


@pytest.mark.gpu
def test_fortran_restart_is_bitwise(tmp_path):
    case = cases.tgv3d(16)
    for fid, name in ((U, "u"), (V, "v"), (W, "w"), (P, "p")):
        np.asfortranarray(case.fields[fid]).ravel(order="F").tofile(tmp_path / f"init_{name}.raw")
    (tmp_path / "data").mkdir()
    exe = _bin("restart_3d")
    subprocess.run([exe, "1"], cwd=tmp_path, check=True, timeout=120)
    subprocess.run([exe, "2"], cwd=tmp_path, check=True, timeout=120)
    a = (tmp_path / "data" / "state_0000010.raw").read_bytes()
    b = (tmp_path / "data" / "state_0000020.raw").read_bytes()
    assert len(a) == 7 * 16 ** 3 * 8 and a == b


@pytest.mark.gpu
def test_fortran_rayleigh_taylor_matches_oracle(tmp_path):
    """The reference's Rayleigh-Taylor program (two-phase build) at 32 x 128,
    30 steps, against the oracle's MF steps from the same initial vof.  exp,
    log, tanh and cosh come from the device's and the host's libm, so the bar
    is the two-phase one: 1e-9 relative."""
    import math
    from fen_amd._lib import BC_PERIODIC, BC_WALL, VOF
    from oracle import oracle as O
    nx, steps = 32, 30
    (tmp_path / "data").mkdir()
    out = subprocess.run([_bin("rayleigh_taylor_2d"), str(steps), str(nx)], cwd=tmp_path, check=True,
                         capture_output=True, text=True, timeout=120).stdout
    status = [l for l in out.splitlines() if l.startswith("step:")]
    assert len(status) == steps
    n = (nx, 4 * nx)
    d = 1.0 / nx
    vof0 = cases.vof_from_distance(lambda x, y: y - 0.05 * np.cos(2.0 * math.pi * x) - 2.0, n[0], n[1], d)
    saved0 = np.fromfile(tmp_path / "data" / "vof_0").reshape(n, order="F")
    assert np.abs(saved0 - vof0).max() < 1e-15
    o = O.Oracle(2, (n[0], n[1], 1), (1.0, 4.0, d),
                 bc=[BC_PERIODIC, BC_PERIODIC, BC_WALL, BC_WALL, BC_PERIODIC, BC_PERIODIC])
    o.set_params(density=1.0, viscosity=1.0, g=(0.0, -9.80665, 0.0), cfl=1.0)
    o.set_multiphase(0.1694, 1.225, 3.13e-3, 3.13e-3, 0.0)
    o.field(VOF)[1:-1, 1:-1, 1] = saved0
    o.update_ghost_nodes(VOF)
    o.update_material_properties()
    dt = o.mf_set_timestep(1.0)
    for s in range(1, steps + 1):
        _, md, _ = o.step(s, dt)
    dt_f = float(status[-1][38:51])
    assert abs(dt_f - dt) <= 1e-6 * dt
    for name, fid in (("u", U), ("v", V), ("p", P), ("vof", VOF)):
        got = np.fromfile(tmp_path / f"{name}.raw").reshape(n, order="F")
        ref = o.field(fid)[1:-1, 1:-1, 1]
        err = rel_err(got, ref)
        print(f"Fortran Rayleigh-Taylor {n[0]}x{n[1]}, {steps} steps: {name} rel err {err:.2e}")
        assert err < 1e-9, name

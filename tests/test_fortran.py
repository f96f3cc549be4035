"""FEN's Fortran interface over libfen: the layer compiles for DIM=2 and
DIM=3 (CPU), and FEN-style programs built on it reproduce the oracle and
the reference's restart criterion on the GPU."""
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
@pytest.mark.parametrize("dim", [2, 3])
def test_fortran_layer_compiles(dim, tmp_path):
    for src in ("fen_capi.f90", "fen_modules.F90"):
        subprocess.run([FC, "-cpp", f"-DDIM={dim}", "-c", os.path.join(FDIR, src)], cwd=tmp_path, check=True)
    assert (tmp_path / "navier_stokes_mod.mod").exists() and (tmp_path / "poisson_mod.mod").exists()


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

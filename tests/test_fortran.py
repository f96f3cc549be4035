"""FEN's Fortran interface over libfen: the layer compiles for DIM=2 and
DIM=3 and in the two-phase build (CPU), and FEN-style programs built on it
reproduce the oracle and the reference's restart criterion on the GPU."""
import json
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


def _case(path, grid, steps, **extra):
    c = {"Grid": grid, "Timestep": extra.pop("timestep", {"U": 2.0}), "Run": {"steps": steps, **extra.pop("run", {})}}
    c.update(extra)
    path.write_text(json.dumps(c, indent=2))
    return path


def _raw(a, path):
    np.asfortranarray(a).ravel(order="F").tofile(path)


@pytest.mark.gpu
def test_fortran_restart_is_bitwise(tmp_path):
    """The reference's restart criterion (test/small_test/io/run.sh:6-16):
    a run that saves its state after steps 5 and 10, and a run restarted
    from step 5 that saves after step 10, write identical state files.
    Both runs are case files driven by run_case."""
    case = cases.tgv3d(16)
    for fid, name in ((U, "u"), (V, "v"), (W, "w"), (P, "p")):
        _raw(case.fields[fid], tmp_path / f"init_{name}.raw")
    grid = {"Nx": 16, "Ny": 16, "Nz": 16, "Lx": 2 * np.pi, "Ly": 2 * np.pi, "Lz": 2 * np.pi}
    (tmp_path / "data").mkdir()
    (tmp_path / "again").mkdir()
    first = _case(tmp_path / "first.json", grid, 10, Initial={k: f"init_{k}.raw" for k in "uvwp"},
                  run={"save_state": [5, 10], "status_every": 0})
    second = _case(tmp_path / "second.json", grid, 5, Restart={"step": 5, "dir": "data"},
                   run={"save_state": [10], "data_dir": "again", "status_every": 0})
    exe = _bin("run_case_3d")
    subprocess.run([exe, str(first)], cwd=tmp_path, check=True, timeout=120)
    subprocess.run([exe, str(second)], cwd=tmp_path, check=True, timeout=120)
    a = (tmp_path / "data" / "state_0000010.raw").read_bytes()
    b = (tmp_path / "again" / "state_0000010.raw").read_bytes()
    assert len(a) == 7 * 16 ** 3 * 8 and a == b


@pytest.mark.gpu
def test_fortran_case_file_run_matches_oracle(tmp_path):
    """A 3-D case file through run_case (set-up, status lines, save_fields,
    final fields) against the oracle; grid%setup's grid.json equals the
    C-ABI's fen_write_grid_json for the same grid."""
    from fen_amd import Solver
    N, steps = 32, 12
    case = cases.tgv3d(N, steps=steps)
    for fid, name in ((U, "u"), (V, "v"), (W, "w"), (P, "p")):
        _raw(case.fields[fid], tmp_path / f"u0_{name}.raw")
    L = 2 * np.pi
    grid = {"Nx": N, "Ny": N, "Nz": N, "Lx": L, "Ly": L, "Lz": L, "name": "tgv"}
    (tmp_path / "out").mkdir()
    path = _case(tmp_path / "case.json", grid, steps, timestep={"dt": case.dt},
                 Physics={"density": case.density, "viscosity": case.viscosity, "CFL": case.cfl},
                 Initial={k: f"u0_{k}.raw" for k in "uvwp"},
                 run={"status_every": 4, "save_fields": [steps], "data_dir": "out", "write_final": True})
    out = subprocess.run([_bin("run_case_3d"), str(path)], cwd=tmp_path, check=True, capture_output=True,
                         text=True, timeout=120).stdout
    status = [l for l in out.splitlines() if l.startswith("step:")]
    assert [int(l[6:13]) for l in status] == [4, 8, 12] and all(len(l) == 97 for l in status)
    o = make_oracle(case)
    for s in range(1, steps + 1):
        _, md, mc = o.step(s, case.dt)
    # maxdiv is round-off on both sides; maxCFL to the 7 digits of E13.6
    assert abs(float(status[-1][60:73])) < 1e-13 and abs(md) < 1e-13
    assert abs(float(status[-1][84:97]) - mc) <= 1e-5 * mc   # E13.6: six significant digits
    for name, fid in (("u", U), ("v", V), ("w", W), ("p", P)):
        got = np.fromfile(tmp_path / f"{name}.raw").reshape((N, N, N), order="F")
        assert rel_err(got, o.field(fid)[1:-1, 1:-1, 1:-1]) < 1e-12, name
    assert (tmp_path / "out" / f"vz_{steps:07d}.raw").stat().st_size == N ** 3 * 8
    # grid%setup wrote <name>.json (src/grid.f90:198)
    g = Solver(3, (N, N, N), (L, L, L))
    g.write_grid_json(str(tmp_path / "c.json"))
    assert (tmp_path / "c.json").read_bytes() == (tmp_path / "tgv.json").read_bytes()
    g.close()


@pytest.mark.gpu
def test_fortran_rayleigh_taylor_matches_oracle(tmp_path):
    """A two-phase Rayleigh-Taylor case file (32 x 128 on [0,1] x [0,4],
    periodic in x, walls in y, heavy fluid 1.225 above light 0.1694 across
    y = 2 + 0.05 cos(2 pi x), mu = 3.13e-3, gravity -9.80665, dt from
    set_timestep with U = 1) run by run_case in the two-phase build, 30
    steps, against the oracle's MF steps from the same initial vof.  exp,
    log, tanh and cosh come from the device's and the host's libm, so the
    bar is the two-phase one: 1e-9 relative."""
    import math
    from fen_amd._lib import BC_PERIODIC, BC_WALL, VOF
    from oracle import oracle as O
    nx, steps = 32, 30
    n = (nx, 4 * nx)
    d = 1.0 / nx
    grid = {"Nx": nx, "Ny": 4 * nx, "Nz": 1, "Lx": 1.0, "Ly": 4.0,
            "boundary_conditions": ["Periodic", "Periodic", "Wall", "Wall"]}
    path = _case(tmp_path / "rt.json", grid, steps, timestep={"U": 1.0},
                 Physics={"viscosity": 1.0, "g": [0.0, -9.80665, 0.0]},
                 Multiphase={"rho_0": 0.1694, "rho_1": 1.225, "mu_0": 3.13e-3, "mu_1": 3.13e-3,
                             "interface": {"level": 2.0, "amplitude": 0.05, "wavelength": 1.0}},
                 run={"write_final": True})
    out = subprocess.run([_bin("run_case_2d_mf"), str(path)], cwd=tmp_path, check=True,
                         capture_output=True, text=True, timeout=120).stdout
    status = [l for l in out.splitlines() if l.startswith("step:")]
    assert len(status) == steps
    vof0 = cases.vof_from_distance(lambda x, y: y - 0.05 * np.cos(2.0 * math.pi * x) - 2.0, n[0], n[1], d)
    o = O.Oracle(2, (n[0], n[1], 1), (1.0, 4.0, d),
                 bc=[BC_PERIODIC, BC_PERIODIC, BC_WALL, BC_WALL, BC_PERIODIC, BC_PERIODIC])
    o.set_params(density=1.0, viscosity=1.0, g=(0.0, -9.80665, 0.0), cfl=1.0)
    o.set_multiphase(0.1694, 1.225, 3.13e-3, 3.13e-3, 0.0)
    o.field(VOF)[1:-1, 1:-1, 1] = vof0
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


@pytest.mark.gpu
def test_fortran_two_phase_restart_is_bitwise(tmp_path):
    """The two-phase restart criterion (test/small_test/io/test_MF.f90 with
    run.sh): 10 steps saving the state after steps 5 and 10, and a run
    restarted from step 5 that saves after step 10, write identical files
    (p, u, v, dv_o, vof, rho, mu, p_o).  The file does not hold the vof's
    sweep order; the restarted case sets it from the step count."""
    nx = 16
    grid = {"Nx": nx, "Ny": 4 * nx, "Nz": 1, "Lx": 1.0, "Ly": 4.0,
            "boundary_conditions": ["Periodic", "Periodic", "Wall", "Wall"]}
    common = dict(timestep={"U": 1.0}, Physics={"viscosity": 1.0, "g": [0.0, -9.80665, 0.0]},
                  Multiphase={"rho_0": 0.1694, "rho_1": 1.225, "mu_0": 3.13e-3, "mu_1": 3.13e-3,
                              "interface": {"level": 2.0, "amplitude": 0.05, "wavelength": 1.0}})
    (tmp_path / "data").mkdir()
    (tmp_path / "again").mkdir()
    first = _case(tmp_path / "first.json", grid, 10, run={"save_state": [5, 10], "status_every": 0}, **common)
    second = _case(tmp_path / "second.json", grid, 5, Restart={"step": 5, "dir": "data"},
                   run={"save_state": [10], "data_dir": "again", "status_every": 0}, **common)
    exe = _bin("run_case_2d_mf")
    subprocess.run([exe, str(first)], cwd=tmp_path, check=True, timeout=120)
    subprocess.run([exe, str(second)], cwd=tmp_path, check=True, timeout=120)
    a = (tmp_path / "data" / "state_0000010.raw").read_bytes()
    b = (tmp_path / "again" / "state_0000010.raw").read_bytes()
    assert len(a) == 9 * nx * 4 * nx * 8 and a == b


@pytest.mark.gpu
def test_fortran_stops_on_nonfinite(tmp_path):
    """A NaN in the initial velocity reaches the first step's checks; the
    Fortran layer maps FEN_ERR_NONFINITE to the library's message on stderr
    and an error stop (a non-zero exit), as its error path does for every
    failed call."""
    import math
    N = 16
    case = cases.tgv3d(N, steps=2)
    u = case.fields[U].copy(order="F")
    u[5, 6, 7] = math.nan
    for name, arr in (("u", u), ("v", case.fields[V]), ("w", case.fields[W]), ("p", case.fields[P])):
        _raw(arr, tmp_path / f"u0_{name}.raw")
    L = 2 * np.pi
    path = _case(tmp_path / "case.json", {"Nx": N, "Ny": N, "Nz": N, "Lx": L, "Ly": L, "Lz": L}, 2,
                 timestep={"dt": case.dt}, Initial={k: f"u0_{k}.raw" for k in "uvwp"})
    r = subprocess.run([_bin("run_case_3d"), str(path)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "navier_stokes_solver failed with status -4" in r.stderr, r.stderr
    assert "non-finite velocity or divergence" in r.stderr, r.stderr
    assert not [l for l in r.stdout.splitlines() if l.startswith("step:")]   # no status line was printed

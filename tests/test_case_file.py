"""The case file reader (fen_case_read) and the context-free text formats
(fen_status_text, fen_grid_json_text), on the CPU.

The texts are compared byte for byte with what the Fortran runtime writes
for the reference's own FORMAT statements (src/grid.f90:246-258,
src/navier_stokes.f90:746), compiled with amdflang; the case structure's
layout is compared with the C compiler's and the Fortran binding's.
"""
import ctypes
import json
import os
import subprocess

import pytest

from fen_amd import _lib, case_file
from fen_amd._lib import BC_INFLOW, BC_OUTFLOW, BC_PERIODIC, BC_WALL, FP32_REFINE, FP64

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "fen_amd", "fortran")
FC = "/opt/rocm/bin/amdflang"
needs_fc = pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not installed")


def _write(tmp_path, obj, name="case.json"):
    p = tmp_path / name
    p.write_text(obj if isinstance(obj, str) else json.dumps(obj, indent=2))
    return p


FULL = {
    "Grid": {"Nx": 64, "Ny": 32, "Nz": 16, "origin": [0.5, -1.0, 2.0], "Lx": 4.0, "Ly": 2.0, "Lz": 1.0,
             "boundary_conditions": ["Inflow", "Outflow", "Periodic", "Periodic", "Wall", "Wall"],
             "name": "duct"},
    "Physics": {"density": 2.0, "viscosity": 0.001, "g": [1.0, 0.0, -9.81], "CFL": 0.4, "constant_CFL": True},
    "Timestep": {"U": 1.5},
    "Precision": "FP32_REFINE",
    "Initial": {"u": "u0.raw", "w": "w0.raw", "p": "p0.raw"},
    "Restart": {"step": 40, "dir": "old"},
    "Run": {"steps": 200, "status_every": 10, "save_state": [100, 200], "save_fields": [50, 150, 200],
            "data_dir": "out", "write_final": True},
}


def test_full_case(tmp_path):
    cs = case_file.read(_write(tmp_path, FULL))
    assert cs.dim == 3 and list(cs.n) == [64, 32, 16]
    assert list(cs.origin) == [0.5, -1.0, 2.0] and list(cs.L) == [4.0, 2.0, 1.0]
    assert list(cs.bc) == [BC_INFLOW, BC_OUTFLOW, BC_PERIODIC, BC_PERIODIC, BC_WALL, BC_WALL]
    assert cs.name == b"duct"
    assert (cs.density, cs.viscosity, list(cs.g), cs.cfl, cs.constant_cfl) == (2.0, 0.001, [1.0, 0.0, -9.81], 0.4, 1)
    assert cs.U == 1.5 and cs.dt == 0.0 and cs.precision == FP32_REFINE
    assert [cs.init_file(w) for w in ("u", "v", "w", "p", "vof")] == ["u0.raw", "", "w0.raw", "p0.raw", ""]
    assert cs.restart == 40 and cs.restart_dir == b"old"
    assert (cs.steps, cs.status_every, cs.write_final) == (200, 10, 1)
    assert list(cs.save_state[:cs.n_save_state]) == [100, 200]
    assert list(cs.save_fields[:cs.n_save_fields]) == [50, 150, 200]
    assert cs.data_dir == b"out" and cs.multiphase == 0


def test_defaults_are_fens(tmp_path):
    """Omitted physics take navier_stokes_mod's defaults
    (src/navier_stokes.f90:17-23,42), omitted faces grid%setup's Periodic
    (src/grid.f90:95-104); Nz = 1 is a 2-D case with Lz = Lx Nz / Nx."""
    cs = case_file.read(_write(tmp_path, {"Grid": {"Nx": 128, "Ny": 128, "Nz": 1, "Lx": 6.0, "Ly": 6.0},
                                          "Timestep": {"dt": 1e-3}, "Run": {"steps": 10}}))
    assert cs.dim == 2 and cs.L[2] == 6.0 / 128
    assert (cs.density, cs.viscosity, list(cs.g), cs.cfl, cs.constant_cfl) == (1.0, 1.0, [0.0, 0.0, 0.0], 1.0, 0)
    assert list(cs.bc) == [BC_PERIODIC] * 6 and list(cs.origin) == [0.0, 0.0, 0.0]
    assert cs.precision == FP64 and cs.restart == -1 and cs.dt == 1e-3 and cs.U == 0.0
    assert cs.status_every == 1 and cs.n_save_state == 0 and cs.data_dir == b"data" and cs.name == b"grid"


def test_numbers_are_correctly_rounded(tmp_path):
    text = ('{"Grid": {"Nx": 8, "Ny": 8, "Nz": 8, "Lx": 6.283185307179586, "Ly": 6.283185307179586e0, '
            '"Lz": 0.6283185307179586E+1}, "Timestep": {"dt": 1.0e-3}, "Run": {"steps": 1}}')
    cs = case_file.read(_write(tmp_path, text))
    assert list(cs.L) == [6.283185307179586] * 3 and cs.dt == 0.001


@pytest.mark.parametrize("mutate,msg", [
    (lambda c: c.pop("Grid"), 'missing object "Grid"'),
    (lambda c: c["Grid"].pop("Nx"), 'missing "Nx"'),
    (lambda c: c["Grid"].__setitem__("Nx", 16.5), '"Nx" must be an integer'),
    (lambda c: c["Grid"].__setitem__("Lx", -1.0), "Lx, Ly and Lz must be > 0"),
    (lambda c: c["Grid"].__setitem__("boundary_conditions", ["Wall"] * 4), "must list 6 faces"),
    (lambda c: c["Grid"].__setitem__("boundary_conditions", ["Wall"] * 5 + ["Slip"]), '"Slip" is not Periodic'),
    (lambda c: c["Grid"].__setitem__("Nxx", 1), 'unknown key "Nxx"'),
    (lambda c: c["Timestep"].__setitem__("dt", 1e-3), 'exactly one of "dt"'),
    (lambda c: c.__setitem__("Timestep", {"dt": 1e-3}), 'constant_CFL starts from set_timestep'),
    (lambda c: c["Physics"].__setitem__("density", 0), "density must be > 0"),
    (lambda c: c["Physics"].__setitem__("constant_CFL", 1), "must be true or false"),
    (lambda c: c.__setitem__("Precision", "FP16"), 'must be "FP64" or "FP32_REFINE"'),
    (lambda c: c["Run"].__setitem__("save_state", list(range(33))), "more than 32 steps"),
    (lambda c: c["Run"].__setitem__("steps", -1), "steps must be >= 0"),
    (lambda c: c["Initial"].__setitem__("vof", "f.raw"), 'a vof needs the "Multiphase" object'),
    (lambda c: c.__setitem__("Multiphase", {"rho_0": 1, "rho_1": 2, "mu_0": 1, "mu_1": 1}), "two-phase mode is 2-D"),
])
def test_invalid_cases_are_rejected(tmp_path, mutate, msg):
    c = json.loads(json.dumps(FULL))
    mutate(c)
    with pytest.raises(case_file.CaseError, match=msg.replace("(", r"\(").replace('"', '"')):
        case_file.read(_write(tmp_path, c))


@pytest.mark.parametrize("text,msg", [
    ('{"Grid": {"Nx": 8,}}', "line 1, column 19: expected a key string"),
    ('{\n  "Grid": {"Nx": 8, "Nx": 9}\n}', 'line 2, column 21: duplicate key "Nx"'),
    ('{"Grid": [1, 2}', "expected ',' or ']'"),
    ('{"Grid": {"Nx": 01}}', "expected ',' or '}'"),
    ('{"a": "\\u0041"}', "unsupported escape"),
    ('{} x', "unexpected text after the top-level value"),
    ('', "unexpected end of file"),
    ('[' * 100 + ']' * 100, "nesting deeper than 64 levels"),
])
def test_malformed_json_names_the_place(tmp_path, text, msg):
    with pytest.raises(case_file.CaseError, match=msg.replace("(", r"\(").replace("[", r"\[").replace("]", r"\]")):
        case_file.read(_write(tmp_path, text))


def test_missing_file(tmp_path):
    with pytest.raises(case_file.CaseError, match="cannot open case file"):
        case_file.read(tmp_path / "nope.json")


def test_multiphase_case(tmp_path):
    c = {"Grid": {"Nx": 32, "Ny": 128, "Nz": 1, "Lx": 1.0, "Ly": 4.0,
                  "boundary_conditions": ["Periodic", "Periodic", "Wall", "Wall"]},
         "Physics": {"viscosity": 1.0, "g": [0.0, -9.80665, 0.0]},
         "Timestep": {"U": 1.0}, "Run": {"steps": 30},
         "Multiphase": {"rho_0": 0.1694, "rho_1": 1.225, "mu_0": 3.13e-3, "mu_1": 3.13e-3,
                        "interface": {"level": 2.0, "amplitude": 0.05}}}
    cs = case_file.read(_write(tmp_path, c))
    assert cs.multiphase == 1 and (cs.rho0, cs.rho1, cs.mu0, cs.mu1, cs.sigma) == (0.1694, 1.225, 3.13e-3, 3.13e-3, 0)
    assert list(cs.iface) == [2.0, 0.05, 1.0]   # wavelength defaults to Lx
    assert list(cs.bc[:4]) == [BC_PERIODIC, BC_PERIODIC, BC_WALL, BC_WALL]


# ------------------------------------------------------------ structure layout
def test_struct_layout_matches_c(tmp_path):
    """sizeof and member offsets of struct fen_case as gcc lays it out."""
    members = [name for name, _ in case_file.FenCase._fields_]
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"fen.h\"\nint main(void){\n"
                   + 'printf("%zu\\n", sizeof(fen_case));\n'
                   + "".join(f'printf("%zu\\n", offsetof(fen_case, {m}));\n' for m in members) + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(case_file.FenCase)
    assert vals[1:] == [getattr(case_file.FenCase, m).offset for m in members]


@needs_fc
def test_fortran_binding_reads_the_same_case(tmp_path):
    """The bind(C) type of fen_capi.f90 receives the members fen_case_read
    writes (a size mismatch or a shifted member shows here)."""
    subprocess.run([FC, "-cpp", "-DDIM=3", "-c", os.path.join(FDIR, "fen_capi.f90")], cwd=tmp_path, check=True)
    prog = tmp_path / "rd.f90"
    prog.write_text("""program rd
    use iso_c_binding
    use fen_capi
    type(fen_case) :: cs
    character(kind=c_char) :: err(256)
    integer(c_int) :: rc
    rc = fen_case_read('case.json'//c_null_char, cs, err, 256_c_size_t)
    write(*, '(I0)') c_sizeof(cs)
    write(*, '(I0,1x,I0,1x,I0,1x,I0)') rc, cs%n(1), cs%bc(6), cs%steps
    write(*, '(I0,1x,I0,1x,I0)') cs%n_save_fields, cs%save_fields(3), cs%write_final
    write(*, '(ES24.16,1x,ES24.16,1x,ES24.16)') cs%origin(1), cs%g(3), cs%U
    write(*, '(6A)') cs%init(1:6, 4)
    write(*, '(3A)') cs%data_dir(1:3)
end program rd
""")
    lib = os.path.join(ROOT, "fen_amd")
    subprocess.run([FC, "rd.f90", "fen_capi.o", "-o", "rd", f"-L{lib}", "-lfen", f"-Wl,-rpath,{lib}"],
                   cwd=tmp_path, check=True)
    _write(tmp_path, FULL)
    out = subprocess.run(["./rd"], cwd=tmp_path, capture_output=True, text=True, check=True).stdout.split("\n")
    assert int(out[0]) == ctypes.sizeof(case_file.FenCase)
    assert out[1].split() == ["0", "64", str(BC_WALL), "200"]
    assert out[2].split() == ["3", "200", "1"]
    assert [float(x) for x in out[3].split()] == [0.5, -9.81, 1.5]
    assert out[4] == "p0.raw" and out[5] == "out"


# ------------------------------------------------------- text formats vs Fortran
STATUS_VALUES = [
    (1, 1e-3, 1e-3, 1.2345678e-15, 0.5),
    (1000, 1.0, 1e-3, -3.0e-17, 0.123456789),
    (200, 1.2271846303085130, 6.135923151542565e-3, 0.0, 1.9999999),
    (9999999, 12345.678, 1e-120, -1e-300, 9.99999951e99),
    (42, -0.5, 2.5e-5, 6.484396353201305e-15, 0.0),
]

GRIDS = [
    (256, 256, 256, (0.0, 0.0, 0.0), (6.283185307179586,) * 3),
    (512, 512, 256, (0.0, 0.0, 0.0), (4.0, 4.0, 2.0)),
    (128, 128, 1, (-1.5, 0.25, 0.0), (6.283185307179586, 6.283185307179586, 6.283185307179586 / 128)),
    (32, 128, 1, (1e-9, -123.456, 0.0), (1.0, 4.0, 1.0 / 32)),
]


@needs_fc
def test_status_text_matches_fortran_format(tmp_path):
    lines = []
    for step, time, dt, md, mc in STATUS_VALUES:
        lines.append(f"    write(*, 10) 'step: ', {step}, 'time: ', {time!r}_8, 'dt: ', {dt!r}_8, "
                     f"'maxdiv: ', {md!r}_8, 'maxCFL: ', {mc!r}_8")
    prog = tmp_path / "st.f90"
    prog.write_text("program st\n" + "\n".join(lines) +
                    "\n10  format(A6,I7,1x,A6,E13.6,1x,A4,E13.6,1x,A8,E13.6,1x,A9,1x,E13.6)\nend program st\n")
    subprocess.run([FC, "st.f90", "-o", "st"], cwd=tmp_path, check=True)
    want = subprocess.run(["./st"], cwd=tmp_path, capture_output=True, text=True, check=True).stdout.splitlines()
    lib = _lib.load()
    buf = ctypes.create_string_buffer(256)
    for v, w in zip(STATUS_VALUES, want):
        assert lib.fen_status_text(*v, buf, len(buf)) == 0
        assert buf.value.decode() == w
    assert lib.fen_status_text(1, 0.0, 0.0, 0.0, 0.0, buf, 20) == -1   # too small: refused


@needs_fc
def test_grid_json_matches_fortran_print_json(tmp_path):
    """grid%print_json of the Fortran layer (run by grid%setup, as in the
    reference) against fen_grid_json_text, byte for byte."""
    for dim in (2, 3):
        for src in ("fen_capi.f90", "fen_modules.F90"):
            subprocess.run([FC, "-cpp", f"-DDIM={dim}", "-c", os.path.join(FDIR, src)], cwd=tmp_path, check=True)
        prog = tmp_path / "gj.f90"
        prog.write_text("""program gj
    use precision_mod, only : dp
    use grid_mod
    type(grid) :: g
    integer :: n(3)
    real(dp) :: x0(3), L(3)
    read(*, *) n, x0, L
    call g%setup(n(1), n(2), n(3), L(1), L(2), L(3), x0, 1, 1)
end program gj
""")
        lib_dir = os.path.join(ROOT, "fen_amd")
        subprocess.run([FC, "-cpp", f"-DDIM={dim}", "gj.f90", "fen_modules.o", "fen_capi.o", "-o", f"gj{dim}",
                        f"-L{lib_dir}", "-lfen", f"-Wl,-rpath,{lib_dir}"], cwd=tmp_path, check=True)
        lib = _lib.load()
        for nx, ny, nz, x0, L in GRIDS:
            if (nz == 1) != (dim == 2):
                continue
            inp = f"{nx} {ny} {nz} {x0[0]!r} {x0[1]!r} {x0[2]!r} {L[0]!r} {L[1]!r} {L[2]!r}\n"
            subprocess.run([f"./gj{dim}"], cwd=tmp_path, input=inp, text=True, check=True)
            want = (tmp_path / "grid.json").read_bytes()
            buf = ctypes.create_string_buffer(1024)
            x0a = (ctypes.c_double * 3)(*x0)
            assert lib.fen_grid_json_text(nx, ny, nz, x0a, *L, buf, len(buf)) == 0
            assert buf.value == want, (nx, ny, nz)
            # the written grid.json is JSON, and its Grid object is a case file's
            assert json.loads(want)["Grid"]["Nx"] == nx

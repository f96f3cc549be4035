"""Field output (fen_amd.output): save_fields' cell-centred raw files and the
XDMF index of a run directory, on host arrays; one GPU test writes a run
directory from a resident state."""
import json
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from fen_amd import output
from fen_amd._lib import P, U, V, W


def test_face_to_center_matches_the_reference_average():
    """sc(i,j,k) = 0.5 (sf(i,j,k) + sf(i-di,j-dj,k-dk)) over the interior (src/fields.f90:210)."""
    rng = np.random.default_rng(2)
    f = rng.standard_normal((7, 6, 5))
    for axis, (di, dj, dk) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
        c = output.face_to_center(f, axis)
        assert c.shape == (5, 4, 3)
        for i in range(1, 6):
            for j in range(1, 5):
                for k in range(1, 4):
                    assert c[i - 1, j - 1, k - 1] == 0.5 * (f[i, j, k] + f[i - di, j - dj, k - dk])


class _FakeSolver:
    def __init__(self, dim, fields):
        self.dim, self.fields = dim, fields

    def download(self, fid):
        return self.fields[fid]


def test_save_fields_writes_fortran_ordered_interiors(tmp_path):
    rng = np.random.default_rng(3)
    shp = (6, 5, 4)
    fields = {fid: np.asfortranarray(rng.standard_normal(shp)) for fid in (U, V, W, P)}
    paths = output.save_fields(_FakeSolver(3, fields), 40, str(tmp_path / "data"))
    assert [p.split("/")[-1] for p in paths] == ["vx_0000040.raw", "vy_0000040.raw", "vz_0000040.raw",
                                                 "p_0000040.raw"]
    n = (4, 3, 2)
    for path, fid, axis in zip(paths, (U, V, W), (0, 1, 2)):
        got = np.fromfile(path, dtype="<f8").reshape(n, order="F")
        assert np.array_equal(got, output.face_to_center(fields[fid], axis))
    got = np.fromfile(paths[3], dtype="<f8").reshape(n, order="F")
    assert np.array_equal(got, fields[P][1:-1, 1:-1, 1:-1])


def _grid_json(path, n, L, x0=(0.0, 0.0, 0.0)):
    """grid.json in the layout print_json writes (src/grid.f90)."""
    e = lambda v: f"{v:.8E}"
    text = ("{\n    \"Grid\": {\n"
            f"        \"Nx\":  {n[0]:7d},\n        \"Ny\":  {n[1]:7d},\n        \"Nz\":  {n[2]:7d},\n"
            f"        \"origin\": [ {e(x0[0])},{e(x0[1])},{e(x0[2])}],\n"
            f"        \"Lx\":  {e(L[0])},\n        \"Ly\":  {e(L[1])},\n        \"Lz\":  {e(L[2])}\n"
            "    }\n}\n")
    json.loads(text)
    path.write_text(text)


def test_write_xdmf_indexes_saved_steps(tmp_path):
    n, L = (8, 4, 2), (2.0, 1.0, 0.5)
    _grid_json(tmp_path / "grid.json", n, L, x0=(1.0, 0.0, -0.5))
    data = tmp_path / "data"
    data.mkdir()
    for step in (0, 25, 50):
        for prefix in ("vx", "vy", "vz", "p"):
            np.zeros(n).tofile(data / f"{prefix}_{step:07d}.raw")
    path = output.write_xdmf(str(tmp_path), 25)
    assert path.endswith("3Dfields.xmf")
    root = ET.parse(path).getroot()
    grids = root.find("Domain").find("Grid").findall("Grid")
    assert len(grids) == 3
    last = grids[2]
    assert last.find("Topology").get("Dimensions") == "3 5 9"
    attrs = {a.get("Name"): a.find("DataItem") for a in last.findall("Attribute")}
    assert sorted(attrs) == ["P", "U", "V", "W"]
    assert attrs["W"].text.strip() == "data/vz_0000050.raw"
    assert attrs["U"].get("Dimensions") == "2 4 8"
    geo = [d.text.split() for d in last.find("Geometry").findall("DataItem")]
    assert [float(v) for v in geo[0]] == [-0.5, 0.0, 1.0]          # origin z, y, x
    assert [float(v) for v in geo[1]] == [0.25, 0.25, 0.25]        # spacing


def test_write_xdmf_two_dimensional(tmp_path):
    n, L = (16, 8, 1), (2.0, 1.0, 0.125)
    _grid_json(tmp_path / "grid.json", n, L)
    data = tmp_path / "data"
    data.mkdir()
    for prefix in ("vx", "vy", "p"):
        np.zeros((16, 8)).tofile(data / f"{prefix}_0000000.raw")
    root = ET.parse(output.write_xdmf(str(tmp_path), 10)).getroot()
    g = root.find("Domain").find("Grid").findall("Grid")
    assert len(g) == 1
    assert g[0].find("Topology").get("TopologyType") == "2DCORECTMesh"
    assert sorted(a.get("Name") for a in g[0].findall("Attribute")) == ["P", "U", "V"]


@pytest.mark.gpu
def test_run_directory_output_on_gpu(tmp_path):
    """A run writes grid.json and saves fields every 2 steps; the XDMF index
    and the raw files agree with the resident state."""
    from fen_amd import cases
    from tests.common import make_gpu
    case = cases.tgv3d(16, steps=4)
    g = make_gpu(case)
    g.write_grid_json(str(tmp_path / "grid.json"))
    data = str(tmp_path / "data")
    output.save_fields(g, 0, data)
    g.run(0, 2, case.dt)
    paths = output.save_fields(g, 2, data)
    got = np.fromfile(paths[0], dtype="<f8").reshape(case.n, order="F")
    assert np.array_equal(got, output.face_to_center(g.download(U), 0))
    root = ET.parse(output.write_xdmf(str(tmp_path), 2)).getroot()
    grids = root.find("Domain").find("Grid").findall("Grid")
    assert len(grids) == 2
    assert grids[1].find("Topology").get("Dimensions") == "17 17 17"
    g.close()

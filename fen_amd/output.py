"""Field output of a run, in the reference's file formats.

* ``save_fields(solver, step, directory)`` writes the cell-centred velocity
  components and the pressure as raw float64 interiors in Fortran order,
  ``vx_<step>.raw``, ``vy_...``, ``vz_...`` (3-D) and ``p_...`` with the step
  zero-padded to 7 digits, as ``save_fields`` does (reference
  src/solver.f90:103-156; the face-to-centre average is src/fields.f90:210).
* ``write_xdmf(run_dir, iout)`` indexes the saved steps 0, iout, 2 iout, ...
  in an XDMF file (``2Dfields.xmf`` / ``3Dfields.xmf``) for ParaView or VisIt,
  as the reference's post-processing scripts do (scripts/write2DXDMF.py,
  scripts/write3DXDMF.py): a temporal collection of uniform grids with
  cell-centred U, V, (W,) P attributes read from the raw files.  The grid
  comes from ``grid.json`` (``Solver.write_grid_json``, reference
  src/grid.f90 print_json).

Command line, in a run directory: ``python -m fen_amd.output xdmf IOUT``.
"""
from __future__ import annotations

import argparse
import glob
import json
import os

import numpy as np

from ._lib import P, U, V, W


def face_to_center(f, axis):
    """0.5 (f(i) + f(i-1)) along `axis` over the interior of a ghosted array."""
    nx, ny, nz = (s - 2 for s in f.shape)
    lo = [1, 1, 1]
    lo[axis] = 0
    here = f[1:nx + 1, 1:ny + 1, 1:nz + 1]
    below = f[lo[0]:lo[0] + nx, lo[1]:lo[1] + ny, lo[2]:lo[2] + nz]
    return 0.5 * (here + below)


def _write_raw(path, interior):
    np.asarray(interior, dtype=np.float64).ravel(order="F").tofile(path)   # Fortran order: i fastest


def save_fields(solver, step, directory="data"):
    """Cell-centred u, v, (w) and p of the solver's current state, as raw files."""
    os.makedirs(directory, exist_ok=True)
    sn = f"{int(step):07d}"
    names = [("vx", U, 0), ("vy", V, 1)] + ([("vz", W, 2)] if solver.dim == 3 else [])
    paths = []
    for name, fid, axis in names:
        path = os.path.join(directory, f"{name}_{sn}.raw")
        _write_raw(path, face_to_center(solver.download(fid), axis))
        paths.append(path)
    p = solver.download(P)
    path = os.path.join(directory, f"p_{sn}.raw")
    _write_raw(path, p[1:-1, 1:-1, 1:-1])
    paths.append(path)
    return paths


def _attribute(name, dims, path):
    return (f'    <Attribute Name="{name}" AttributeType="Scalar" Center="Cell">\n'
            f'      <DataItem Dimensions="{dims}" NumberType="Float" Precision="8" Endian="Little" '
            f'Format="Binary">\n        {path}\n      </DataItem>\n    </Attribute>\n')


def write_xdmf(run_dir=".", iout=1, data_dir="data"):
    """XDMF index of the fields saved every `iout` steps under run_dir/data_dir.
    Returns the path of the written file."""
    grid = json.load(open(os.path.join(run_dir, "grid.json")))["Grid"]
    nx, ny, nz = grid["Nx"], grid["Ny"], grid["Nz"]
    lx, ly, lz = grid["Lx"], grid["Ly"], grid["Lz"]
    x0 = grid.get("origin", [0.0, 0.0, 0.0])
    dx, dy, dz = lx / nx, ly / ny, lz / nz
    three_d = os.path.exists(os.path.join(run_dir, data_dir)) and bool(
        glob.glob(os.path.join(run_dir, data_dir, "vz_*")))
    nsaved = len(glob.glob(os.path.join(run_dir, data_dir, "vx_*")))
    out = [
        '<?xml version="1.0" ?>\n<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>\n'
        '<Xdmf xmlns:xi="http://www.w3.org/2003/XInclude" Version="2.2">\n<Domain>\n'
        '<Grid Name="Box" GridType="Collection" CollectionType="Temporal">\n']
    for n in range(nsaved):
        sn = f"{n * iout:07d}"
        if three_d:
            topo = f'<Topology TopologyType="3DCORECTMesh" Dimensions="{nz + 1} {ny + 1} {nx + 1}"/>'
            origin, spacing = f"{x0[2]:16.8e} {x0[1]:16.8e} {x0[0]:16.8e}", f"{dz:16.8e} {dy:16.8e} {dx:16.8e}"
            dims = f"{nz} {ny} {nx}"
            fields = [("U", "vx"), ("V", "vy"), ("W", "vz"), ("P", "p")]
        else:
            topo = f'<Topology TopologyType="2DCORECTMesh" Dimensions="1 {ny + 1} {nx + 1}"/>'
            origin, spacing = "0.0 0.0 0.0", f"{dx:16.8e} {dx:16.8e} {dy:16.8e}"
            dims = f"{ny} {nx}"
            fields = [("U", "vx"), ("V", "vy"), ("P", "p")]
        out.append(
            f'  <Grid Name="Box {n}" GridType="Uniform">\n    {topo}\n'
            '    <Geometry GeometryType="ORIGIN_DXDYDZ">\n'
            f'      <DataItem Name="Origin" Dimensions="3" NumberType="Float" Precision="4" Format="XML">\n'
            f'        {origin}\n      </DataItem>\n'
            f'      <DataItem Name="Spacing" Dimensions="3" NumberType="Float" Precision="4" Format="XML">\n'
            f'        {spacing}\n      </DataItem>\n    </Geometry>\n    <Time Value="{n}" />\n')
        for name, prefix in fields:
            out.append(_attribute(name, dims, f"{data_dir}/{prefix}_{sn}.raw"))
        out.append("  </Grid>\n")
    out.append("</Grid>\n</Domain>\n</Xdmf>\n")
    path = os.path.join(run_dir, "3Dfields.xmf" if three_d else "2Dfields.xmf")
    with open(path, "w") as f:
        f.write("".join(out))
    return path


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    x = sub.add_parser("xdmf", help="index saved fields in an XDMF file")
    x.add_argument("iout", type=int, help="steps between saved fields")
    x.add_argument("--dir", default=".", help="run directory holding grid.json and data/")
    args = ap.parse_args(argv)
    print(write_xdmf(args.dir, args.iout))


if __name__ == "__main__":
    main()

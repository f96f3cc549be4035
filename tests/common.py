"""Helpers shared by the tests: matched oracle / GPU instances of one case."""
from __future__ import annotations

import numpy as np

from fen_amd import cases
from fen_amd._lib import P, PHI, U, V, W
from oracle import oracle as O


def ghosted(interior, dim):
    nx, ny, nz = interior.shape
    g = np.zeros((nx + 2, ny + 2, nz + 2), order="F")
    g[1:-1, 1:-1, 1:-1] = interior
    return g


def make_oracle(case):
    o = O.Oracle(case.dim, case.n, case.L, bc=case.bc)

    def up(fid, arr):
        f = o.field(fid)
        if fid <= W:
            f[1:-1, 1:-1, 1:-1] = arr
        else:
            f[...] = arr

    case.apply(o, up)
    return o


def make_gpu(case, precision=0):
    return cases.make_solver(case, precision)


def live(a, dim):
    """The part of a ghosted array the reference keeps meaningful (2-D: k=1 plane)."""
    return a[:, :, 1:2] if dim == 2 else a


def rel_err(a, b):
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)

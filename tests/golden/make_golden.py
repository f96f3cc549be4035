"""Writes the golden fixtures of this directory with the CPU oracle.

The reference ships no golden vectors for the Navier-Stokes / Poisson path,
and it cannot be built here (FFTW3 and 2decomp&FFT are absent), so these
fixtures are oracle outputs: they pin the oracle against regressions and
give the GPU tests a fixed target.  Inputs are the seeded cases of
fen_amd/cases.py.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from fen_amd import cases  # noqa: E402
from fen_amd._lib import P, U, V, W  # noqa: E402
from tests.common import make_oracle  # noqa: E402

FIXTURES = {
    "tgv2d_16_10steps": (lambda: cases.tgv2d(16, steps=10), (U, V, P)),
    "tgv3d_16_5steps": (lambda: cases.tgv3d(16, steps=5), (U, V, W, P)),
    "channel_16x8x8_5steps": (lambda: cases.channel(16, 8, 8, steps=5), (U, V, W, P)),
}


def run(name):
    make, fields = FIXTURES[name]
    case = make()
    o = make_oracle(case)
    out = {}
    for s in range(1, case.steps + 1):
        dt, md, mc = o.step(s, case.dt)
    out["maxcfl"] = np.array(mc)
    for f in fields:
        out[{P: "p", U: "u", V: "v", W: "w"}[f]] = o.field(f).copy()
    return out


if __name__ == "__main__":
    for name in FIXTURES:
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **run(name))
        print("wrote", name)

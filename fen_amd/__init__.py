"""fen_amd: FEN's fractional-step Navier-Stokes time advance on MI355X (gfx950).

The compute path is libfen.so (HIP kernels behind the C-ABI in include/fen.h);
this package only binds it.
"""
from ._lib import FenError, load  # noqa: F401
from .solver import Solver  # noqa: F401

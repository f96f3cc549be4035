"""The C-ABI library loads on a machine without a GPU, exports every entry
point include/fen.h declares, and rejects bad arguments before touching
the device (no compute call is made here)."""
import ctypes

import pytest

from fen_amd import _lib


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _lib.declared_symbols()
    assert len(declared) >= 28
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing


def _create(dim=3, n=(16, 16, 16), bc=(0, 0, 0, 0, 0, 0), precision=0, out=True):
    lib = _lib.load()
    h = ctypes.c_void_p()
    x0 = (ctypes.c_double * 3)(0, 0, 0)
    bca = (ctypes.c_int * 6)(*bc)
    return lib.fen_create(ctypes.byref(h) if out else None, dim, n[0], n[1], n[2], 1.0, 1.0, 1.0, x0, bca, precision)


@pytest.mark.parametrize("kwargs,code", [
    (dict(out=False), -1),                          # no output pointer
    (dict(dim=4), -1),                              # DIM must be 2 or 3
    (dict(bc=(0, 0, 1, 1, 0, 0)), -2),              # (T,F,T): the reference stops (src/poisson.f90:91-94)
    (dict(dim=2, bc=(1, 1, 0, 0, 0, 0)), -2),       # 2-D (F,T): the reference stops (:106-109)
    (dict(n=(12, 16, 16)), -2),                     # FFT lengths are powers of two
    (dict(bc=(1, 1, 1, 1, 1, 1), n=(16, 12, 16)), -2),   # nnn: y is a transform direction
    (dict(dim=2, bc=(0, 0, 1, 1, 0, 0), n=(16, 12, 1)), -2),   # pn: y spectra use power-of-two rows
    (dict(bc=(1, 1, 0, 0, 1, 1), precision=1), -2),     # npn, nnn, pn, nn: FP64 only
    (dict(dim=2, bc=(1, 1, 1, 1, 0, 0), precision=1), -2),
    (dict(precision=7), -2),
])
def test_create_rejects_before_device_work(kwargs, code):
    assert _create(**kwargs) == code


def test_null_context_calls_are_rejected():
    lib = _lib.load()
    assert lib.fen_destroy(None) == -1
    assert lib.fen_last_error(None) == b"null context"
    d = ctypes.c_double()
    assert lib.fen_ns_step(None, 1, ctypes.byref(d), None, None) == -1
    assert lib.fen_synchronize(None) == -1


def test_profile_names():
    lib = _lib.load()
    names = [lib.fen_profile_name(k).decode() for k in range(lib.fen_profile_count())]
    assert names[:3] == ["momentum", "ghost", "div_r2c"] and "project" in names
    assert lib.fen_profile_name(99) == b""

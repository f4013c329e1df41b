"""CPU checks of the C ABI: the library loads, exports every entry point that
include/nsp.h declares, and its host setup (radial tables) equals the
oracle's restatement of init_fd/get_dr0/get_dr1 bit for bit.  No GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest

import nsp
from oracle_lib import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "nsp.h")).read()
    return sorted(set(re.findall(r"^\s*(?:nsp_status|int|void\s*\*|const char\s*\*|nsp_handle\s*\*)\s*(nsp_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 15
    L = ctypes.CDLL(nsp.LIB_PATH)
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(nsp.EXPORTS)


@pytest.mark.parametrize("nr", [8, 16, 32, 48, 64, 96, 160])
def test_product_tables_equal_oracle_tables(nr):
    T = nsp.radial_tables(nr)
    o = Oracle(nr, 4, 5, 2 * np.pi, 100.0, 0.01)
    R = o.tables()
    for key in ("r", "D", "Lu", "Lp", "rdr", "dr0", "dr1"):
        assert np.array_equal(T[key], R[key]), key


def test_invalid_parameters_are_errors_not_exits():
    p = nsp.Params()
    L = nsp.lib()
    assert L.nsp_default_params(ctypes.byref(p), 32, 64, 65, 1.0, 100.0, 0.01) == 0
    h = ctypes.c_void_p()
    for field, bad in [("nt", 63), ("nz", 64), ("nr", 4), ("dt", -1.0), ("precision", 7)]:
        q = nsp.Params()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(p))
        setattr(q, field, bad)
        assert L.nsp_create(ctypes.byref(q), 0, ctypes.byref(h)) == 1, field
        assert L.nsp_last_error()
    assert L.nsp_radial_tables(4, *([None] * 7)) == 1


@pytest.mark.parametrize("nro,nrn", [(8, 8), (48, 64), (64, 48), (32, 96), (96, 160)])
@pytest.mark.parametrize("flags", [nsp.REMESH_REFERENCE, nsp.REMESH_MIRROR])
def test_product_remesh_weights_equal_oracle(nro, nrn, flags):
    """Host setup of nsp_remesh (the radial matrix) equals the oracle's
    restatement of mod_remesh.cu:80-111 bit for bit."""
    import oracle_lib
    assert np.array_equal(nsp.remesh_weights(nro, nrn, flags), oracle_lib.remesh_weights(nro, nrn, flags))


def test_remesh_weights_reject_bad_arguments():
    L = nsp.lib()
    buf = np.zeros(2 * 8 * 8)
    assert L.nsp_remesh_weights(4, 8, 0, nsp._dp(buf)) == 1
    assert L.nsp_remesh_weights(8, 8, 2, nsp._dp(buf)) == 1
    assert L.nsp_remesh_weights(8, 8, 0, None) == 1

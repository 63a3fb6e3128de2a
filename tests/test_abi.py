"""The C ABI library: it loads, exports every entry point include/*.h declares, rejects bad
descriptors before touching a device, and its host-side grid equals the oracle's.  CPU only."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# the restart file's entry points live in the host library over HDF5; every other header's in the
# gfx950 library
HOST_IO_HEADERS = {"saras_restart.h"}


def declared_symbols(io=False):
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        if (os.path.basename(h) in HOST_IO_HEADERS) != io:
            continue
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(sr_[a-z0-9_]+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from saras_amd import _lib
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {s[0] for s in _lib.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_restart_library_exports_every_declared_symbol():
    from saras_amd import restart
    lib = C.CDLL(restart.LIB_PATH)
    names = declared_symbols(io=True)
    assert names and all(hasattr(lib, n) for n in names), names
    assert set(names) <= {s[0] for s in restart.SIGNATURES}


@pytest.mark.parametrize("n,L,gtype,beta,faces", [(33, 1.0, 1, 1.0, 0), (257, 1.0, 2, 1.0, 0), (129, 2.0, 2, 1.7, 0),
                                                  (65, 6.283185307179586, 1, 1.0, 1), (129, 1.0, 2, 1.0, 1)])
def test_product_grid_equals_oracle_grid(n, L, gtype, beta, faces):
    from saras_amd import mg
    a = mg.grid_axis(n, L, gtype, beta, faces=bool(faces))
    b = po.grid_axis(n, L, gtype, beta, kind="colloc" if faces else "staggr")
    for key in ("x", "xi_x", "xixx", "xix2"):
        assert np.allclose(a[key], b[key], rtol=1e-15, atol=1e-300), key


def test_create_rejects_bad_descriptor_without_device():
    from saras_amd import _lib
    lib = _lib.load()
    d = _lib.SrMgDesc()
    d.xInd = d.yInd = d.zInd = 5
    d.vcDepth = 5          # must be < min index
    h = C.c_void_p()
    assert lib.sr_mg_create(C.byref(d), C.byref(h)) == _lib.SR_ERR_ARG
    assert b"vcDepth" in lib.sr_last_error()
    d.vcDepth = 2
    d.xInd = 0
    assert lib.sr_mg_create(C.byref(d), C.byref(h)) == _lib.SR_ERR_ARG
    assert lib.sr_grid_axis(2, 1.0, 1, 1.0, 0, None, None, None, None) == _lib.SR_ERR_ARG


def test_ts_create_rejects_bad_descriptor_without_device():
    from saras_amd import _lib, ts
    lib = _lib.load()
    h = C.c_void_p()
    d = ts.ts_desc(ind=(5, 5, 5))
    d.xInd = 1
    assert lib.sr_ts_create(C.byref(d), C.byref(h)) == _lib.SR_ERR_ARG
    d = ts.ts_desc(ind=(5, 5, 5), Re=-1.0)
    assert lib.sr_ts_create(C.byref(d), C.byref(h)) == _lib.SR_ERR_ARG
    assert b"Re" in lib.sr_last_error()
    d = ts.ts_desc(ind=(5, 5, 5), L=(0.0, 1.0, 1.0))
    assert lib.sr_ts_create(C.byref(d), C.byref(h)) == _lib.SR_ERR_ARG
    assert h.value is None


def test_guess_create_rejects_bad_grid_without_device():
    from saras_amd import _lib
    lib = _lib.load()
    d = _lib.SrGuessDesc()
    d.n[0], d.n[1], d.n[2] = 17, 17, 12
    h = C.c_void_p()
    assert lib.sr_guess_create(C.byref(d), C.byref(h)) == _lib.SR_ERR_ARG
    assert b"2^k + 1" in lib.sr_last_error()
    d.n[2] = 5
    assert lib.sr_guess_create(C.byref(d), C.byref(h)) == _lib.SR_ERR_ARG
    # the least-squares predictor keeps at most SR_GUESS_MAX_HISTORY pairs
    d.n[2] = 17
    for bad in (-1, 9):
        d.history = bad
        assert lib.sr_guess_create(C.byref(d), C.byref(h)) == _lib.SR_ERR_ARG
        assert b"history" in lib.sr_last_error()
    assert h.value is None
    ci, co, lv = C.c_int(), C.c_int(), C.c_int()
    assert lib.sr_guess_layer_info(4, C.byref(ci), C.byref(co), C.byref(lv)) == 0
    assert (ci.value, co.value, lv.value) == (32, 64, 2)


def test_ts_scalar_descriptor_checks_without_device():
    from saras_amd import _lib, ts
    lib = _lib.load()
    h = C.c_void_p()
    d = ts.ts_desc(ind=(4, 4, 4), scalar=True, Ra=0.0)
    assert lib.sr_ts_create(C.byref(d), C.byref(h)) == _lib.SR_ERR_ARG
    assert b"Ra" in lib.sr_last_error()
    d = ts.ts_desc(ind=(4, 4, 4), scalar=True, rbc_type=9)
    assert lib.sr_ts_create(C.byref(d), C.byref(h)) == _lib.SR_ERR_ARG
    assert b"rbcType" in lib.sr_last_error()

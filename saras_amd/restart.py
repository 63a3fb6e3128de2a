"""Restart files of SARAS (include/saras_restart.h, libsaras_io.so over HDF5).

`write(path, time, fields)` / `read(path, fields)` move the cores of padded host arrays in and out of
output/restartFile.h5 as writer::writeRestart and reader::readData do (writer.cc:455-519,
reader.cc:213-296).  `save(hydro, path, time)` and `load(hydro, path)` do it for the fields of a
time-step handle (saras_amd.ts.Hydro): Vx, Vy, Vz, P and, for scalar runs, T (hydro_d3.cc:140-143,
scalar_d3.cc:128-132); `load` then imposes the boundary conditions, as the reference's constructor
does after reading (hydro_d3.cc:116, scalar_d3.cc:102-104), and clears the pressure correction."""
import ctypes as C
import os

import numpy as np

from . import _lib

LIB_PATH = os.path.join(os.path.dirname(_lib.LIB_PATH), "libsaras_io.so")
_cp = C.POINTER(C.c_char_p)
_dpp = C.POINTER(C.POINTER(C.c_double))
SIGNATURES = [
    ("sr_restart_write", C.c_int, [C.c_char_p, C.c_double, C.c_int, _cp, C.POINTER(C.c_int), _dpp]),
    ("sr_restart_read", C.c_int, [C.c_char_p, C.c_int, _cp, C.POINTER(C.c_int), _dpp, C.POINTER(C.c_double)]),
    ("sr_restart_last_error", C.c_char_p, []),
]
_L = None


def load_lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _L = L
    return _L


def _check(code):
    if code != _lib.SR_OK:
        raise _lib.SarasError(code, load_lib().sr_restart_last_error().decode())


def _marshal(fields):
    names = list(fields)
    arrs = [fields[n] for n in names]
    for a in arrs:
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"] and a.ndim == 3
    cn = (C.c_char_p * len(names))(*[n.encode() for n in names])
    dims = (C.c_int * (3 * len(names)))(*[d - 2 for a in arrs for d in a.shape])
    ptrs = (C.POINTER(C.c_double) * len(names))(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
    return names, cn, dims, ptrs


def write(path, time, fields):
    """fields: name -> padded float64 host array (its core is written)."""
    _, cn, dims, ptrs = _marshal(fields)
    _check(load_lib().sr_restart_write(os.fsencode(path), float(time), len(fields), cn, dims, ptrs))


def read(path, fields):
    """Fills the core of each padded array of `fields` (name -> array) from the file; returns Time."""
    _, cn, dims, ptrs = _marshal(fields)
    t = C.c_double()
    _check(load_lib().sr_restart_read(os.fsencode(path), len(fields), cn, dims, ptrs, C.byref(t)))
    return t.value


def _handle_fields(hydro):
    from . import ts
    which = [("Vx", ts.VX), ("Vy", ts.VY), ("Vz", ts.VZ), ("P", ts.P)]
    if hydro.desc.scalar:
        which.append(("T", ts.T))
    return which


def save(hydro, path, time):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    write(path, time, {name: hydro.download(w) for name, w in _handle_fields(hydro)})


def load(hydro, path):
    from . import ts
    arrays = {name: np.zeros(hydro.shape(w)) for name, w in _handle_fields(hydro)}
    time = read(path, arrays)
    for name, w in _handle_fields(hydro):
        hydro.upload(w, arrays[name])
    hydro.upload(ts.PP, np.zeros(hydro.shape(ts.PP)))
    for c in range(3):
        hydro.imposeVBC(c)
    if hydro.desc.scalar:
        hydro.imposeTBCs()
    return time

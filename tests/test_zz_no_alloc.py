"""sr_probe_gather_noalloc (include/saras_probes.h), the allocation-free variant of sr_probe_gather,
called through the C ABI: on two calls in a row, the second with fewer probes and fields, its
values equal the allocating entry point's bit for bit (both copy the same doubles), and a probe
count above SR_PROBE_NOALLOC_MAX is refused."""
import ctypes as C

import numpy as np
import pytest

from saras_amd import _lib

pytestmark = pytest.mark.gpu


def _call(fn, fields, probes):
    nf, npr = len(fields), len(probes)
    ptrs = (C.c_void_p * nf)(*[f[0] for f in fields])
    dims = (C.c_int * (3 * nf))(*[d for f in fields for d in f[1]])
    idx = np.ascontiguousarray(np.array(probes, dtype=np.int32).reshape(-1))
    out = np.full((npr, nf), np.nan)
    code = fn(nf, ptrs, dims, npr, idx.ctypes.data_as(C.POINTER(C.c_int)), out.ctypes.data_as(C.POINTER(C.c_double)))
    return code, out


def _random_probes(rng, n, shapes):
    lo = -1
    hi = [min(s[a] for s in shapes) - 2 for a in range(3)]   # within every field's padded range
    return [tuple(int(rng.integers(lo, hi[a] + 1)) for a in range(3)) for _ in range(n)]


def test_noalloc_gather_matches_allocating_gather_twice():
    L = _lib.load()
    rng = np.random.default_rng(11)
    shapes = [(35, 19, 11), (34, 19, 11), (35, 18, 11), (35, 19, 10)]
    host = [rng.standard_normal(s) for s in shapes]
    dev = []
    for h in host:
        d = _lib.DeviceArray(h.shape)
        d.upload(h)
        dev.append(d)
    fields = [(d.ptr, tuple(x - 2 for x in d.shape)) for d in dev]
    # first call: 3000 probes of four fields; second call: 37 probes of two fields
    for nprobes, nf in ((3000, 4), (37, 2)):
        probes = _random_probes(rng, nprobes, shapes[:nf])
        c1, a = _call(L.sr_probe_gather, fields[:nf], probes)
        c2, b = _call(L.sr_probe_gather_noalloc, fields[:nf], probes)
        assert c1 == 0 and c2 == 0, _lib.load().sr_last_error()
        assert np.array_equal(a, b), (nprobes, nf)
        for n, (i, j, k) in enumerate(probes):
            for f in range(nf):
                assert b[n, f] == host[f][i + 1, j + 1, k + 1]
    too_many = _random_probes(rng, 4097, shapes[:1])
    code, _ = _call(L.sr_probe_gather_noalloc, fields[:1], too_many)
    assert code == _lib.SR_ERR_ARG and b"SR_PROBE_NOALLOC_MAX" in L.sr_last_error()
    for d in dev:
        d.free()

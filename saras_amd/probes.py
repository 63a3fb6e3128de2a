"""Field probes of SARAS (lib/io/probes.cc, parser::parseProbes / testProbes) on the device fields
of a time-step handle.

* ``parse_probes``: the probe list of the "Probes" string, restating parser::parseProbes
  (parser.cc:424-543): sets in square brackets, each with three entries, an entry an integer or a
  NumPy-linspace range ``start:end:count``; the list is the product of the three entries' index
  lists, x slowest.
* ``check_probes``: parser::testProbes (parser.cc:545-570), raising where the reference prints
  and exits.
* ``Probes``: probes::placeProbes and probeData for one process owning the whole domain: the values
  come from one device gather (include/saras_probes.h, sr_probe_gather), and each call appends one
  line per placed probe to output/ProbesData.dat (probes.cc:61-143)."""
import ctypes as C
import math
import os

import numpy as np

from . import _lib as _L


def _cxx_round(x):
    """std::round: halfway cases away from zero"""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def _istream_uint(text):
    """iss >> unsigned int: leading whitespace skipped, digits read up to the first non-digit"""
    t = text.lstrip()
    d = ""
    for ch in t:
        if not ch.isdigit():
            break
        d += ch
    if not d:
        raise ValueError(f"probe index {text!r} is not a number")
    return int(d)


def _istream_int(text):
    t = text.lstrip()
    sign = ""
    if t[:1] in "+-" and t[:1]:
        sign, t = t[0], t[1:]
    return int(sign + str(_istream_uint(t)))


def parse_probes(coords):
    """The probes of a "Probes" string, in the order of parser::parseProbes: (i, j, k) triples."""
    s = coords
    probes = []
    while True:
        if "[" not in s or "]" not in s:
            raise ValueError("Probes: expected sets in square brackets")
        lb, rb = s.index("["), s.index("]")
        one = s[lb + 1:lb + 1 + (rb - 1)]   # substr(find('[') + 1, find(']') - 1)
        error_probe = one
        one += ","
        lists = []
        while True:
            first = len(one) - len(one.lstrip(" "))
            entry = one[first:one.index(",")]
            entry = entry.rstrip(" ")
            one = one[one.index(",") + 1:]
            if ":" not in entry:
                lists.append([_istream_int(entry)])
            else:
                start, rest = entry.split(":", 1)
                end, count = (rest.split(":", 1) + [""])[:2] if ":" in rest else (rest, rest)
                a, b, n = _istream_uint(start), _istream_uint(end), _istream_uint(count)
                if a == b or n == 1:
                    lists.append([a])
                else:
                    inc = (float(b) - float(a)) / (float(n) - 1.0)
                    lists.append([a + int(_cxx_round(float(i) * inc)) for i in range(n)])
            if len(one) < 2:
                break
        s = s[s.index("]") + 1:]
        s = s.lstrip(" ")
        if len(lists) != 3:
            raise ValueError(f"ERROR: Number of indices for the probe(s) [{error_probe}] does not match "
                             "dimensionality of problem")
        for x in lists[0]:
            for y in lists[1]:
                for z in lists[2]:
                    probes.append((x, y, z))
        if len(s) < 3:
            break
    return probes


def check_probes(probes, ind):
    """parser::testProbes: each index within 0 .. 2^ind + 1 of its direction."""
    for pr in probes:
        for a, axis in enumerate("XYZ"):
            if pr[a] < 0 or pr[a] > 2 ** ind[a] + 1:
                raise ValueError(f"ERROR: The {axis} index of the probe {list(pr)} lies outside the bounds of the domain")
    return probes


def place_probes(probes, nodes):
    """probes::placeProbes on a single sub-domain: subarrayStarts = 0 and subarrayEnds = N - 1 along x
    and y (grid.cc:196-214), so a probe is kept when 0 <= i < N_x - 1 and 0 <= j < N_y - 1; z is not
    checked.  The reference's root rank still writes a line for every listed probe, the lines of
    unplaced probes from uninitialised memory; only placed probes are recorded here."""
    return [p for p in probes if 0 <= p[0] < nodes[0] - 1 and 0 <= p[1] < nodes[1] - 1]


def format_line(time, probe, values):
    """probes::probeData (probes.cc:116-135): time std::fixed setw(6) setprecision(4); setw(4) before
    the indices, which pads the x index only (a stream width applies to the next output alone);
    each value setw(16) setprecision(8) (fixed stays set on the stream)."""
    out = f"{time:6.4f}\t{probe[0]:4d}\t{probe[1]}\t{probe[2]}\t{values[0]:16.8f}"
    for v in values[1:]:
        out += f"\t{v:16.8f}"
    return out


def gather(fields, probes):
    """Values of device fields at the probes: fields is a list of (device pointer, core sizes);
    returns an array (nprobes, nfields)."""
    L = _L.load()
    nf, npr = len(fields), len(probes)
    out = np.zeros((npr, nf))
    if npr == 0:
        return out
    ptrs = (C.c_void_p * nf)(*[f[0] for f in fields])
    dims = (C.c_int * (3 * nf))(*[d for f in fields for d in f[1]])
    idx = np.ascontiguousarray(np.array(probes, dtype=np.int32).reshape(-1))
    _L.check(L.sr_probe_gather(nf, ptrs, dims, npr, idx.ctypes.data_as(C.POINTER(C.c_int)),
                               out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


class Probes:
    """Probes of the written fields of a time-step handle (Vx, Vy, Vz, P and, for scalar runs, T:
    the writeFields of hydro_d3.cc:140-143 / scalar_d3.cc:128-132) into output/ProbesData.dat."""

    def __init__(self, hydro, probes, path, restart=False):
        from . import ts
        self.h = hydro
        self.names = ["Vx", "Vy", "Vz", "P"] + (["T"] if hydro.desc.scalar else [])
        self.which = [ts.VX, ts.VY, ts.VZ, ts.P] + ([ts.T] if hydro.desc.scalar else [])
        d = hydro.desc
        nodes = ((1 << d.xInd) + 1, (1 << d.yInd) + 1, (1 << d.zInd) + 1)
        self.probes = place_probes(probes, nodes)
        self.fields = [(hydro.device_ptr(w), tuple(s - 2 for s in hydro.shape(w))) for w in self.which]
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        self.fh = open(path, "a")
        if not restart:
            self.fh.write("#VARIABLES = Time, X, Y, Z, " + ", ".join(self.names) + "\n")
            self.fh.flush()

    def probe_data(self, time):
        vals = gather(self.fields, self.probes)
        for p, v in zip(self.probes, vals):
            self.fh.write(format_line(time, p, v) + "\n")
        self.fh.flush()
        return vals

    def close(self):
        if self.fh:
            self.fh.close()
            self.fh = None

"""Python mirror of the SARAS multigrid interface over the gfx950 C ABI.

``Parser`` carries the multigrid and mesh fields of the reference parser (lib/io/parser.h:65-102),
``Grid`` the node metrics of lib/grid/grid.cc, and ``Multigrid`` the public and protected methods
of ``poisson`` / ``multigrid_d3`` (lib/poisson/poisson.h:54-135, 193-219) under their reference
names.  Every method calls the HIP library; there is no CPU path.

Arrays at this interface are the reference's padded float64 layout: shape (Nx+2, Ny+2, Nz+2),
C order, k fastest, element [0] on an axis is logical index -1.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as _L


@dataclass
class Parser:
    """Subset of lib/io/parser.h used by the pressure solve (same field names)."""
    xInd: int = 5
    yInd: int = 5
    zInd: int = 5
    Lx: float = 1.0
    Ly: float = 1.0
    Lz: float = 1.0
    xGrid: int = 1          # 1 uniform, 2 tanh (parser.cc setGrids: "U"/"D" per direction)
    yGrid: int = 1
    zGrid: int = 1
    betaX: float = 1.0
    betaY: float = 1.0
    betaZ: float = 1.0
    xPer: bool = False
    yPer: bool = False
    zPer: bool = False
    vcDepth: int = 4
    vcCount: int = 25
    preSmooth: int = 2
    postSmooth: int = 2
    gsSmooth: int = 1
    resType: int = 2
    mgTolerance: float = 1.0e-6
    # additions of this implementation
    testPoisson: bool = False
    semantics: int = _L.SR_SEM_REFERENCE
    precision: int = _L.SR_PREC_FP64
    padMode: int = 0
    coarseProject: int = 0
    coarseDirect: int = 0
    innerPreSmooth: int = 0     # FP32 cycle of the mixed precisions (0: preSmooth / postSmooth)
    innerPostSmooth: int = 0
    relTolerance: float = 1.0e-10

    @staticmethod
    def from_strings(mesh="UUU", domain="PPP", **kw):
        """Mesh type ("U" uniform, "D" double-sided tanh) and domain type ("P" periodic, "N" not),
        as the reference parses them (lib/io/parser.cc:363-414)."""
        p = Parser(**kw)
        p.xGrid, p.yGrid, p.zGrid = (1 if c == "U" else 2 for c in mesh)
        p.xPer, p.yPer, p.zPer = (c == "P" for c in domain)
        return p


def grid_axis(n, length=1.0, grid_type=1, beta=1.0, faces=False):
    """Coordinates and metrics of one direction (lib/grid/grid.cc), padded like the reference."""
    lib = _L.load()
    m = n + 1 if faces else n + 2
    x, xi_x, xixx, xix2 = (np.zeros(m) for _ in range(4))
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    _L.check(lib.sr_grid_axis(int(n), float(length), int(grid_type), float(beta), int(bool(faces)),
                              ptr(x), ptr(xi_x), ptr(xixx), ptr(xix2)))
    return {"x": x, "xi_x": xi_x, "xixx": xixx, "xix2": xix2}


class Grid:
    """Node ("staggered") metrics of the three directions, as grid.h names them."""

    def __init__(self, p: Parser):
        self.p = p
        self.n = [(1 << p.xInd) + 1, (1 << p.yInd) + 1, (1 << p.zInd) + 1]
        ax = [grid_axis(self.n[0], p.Lx, p.xGrid, p.betaX),
              grid_axis(self.n[1], p.Ly, p.yGrid, p.betaY),
              grid_axis(self.n[2], p.Lz, p.zGrid, p.betaZ)]
        self.xStaggr, self.yStaggr, self.zStaggr = ax[0]["x"], ax[1]["x"], ax[2]["x"]
        self.xix2Staggr, self.xixxStaggr = ax[0]["xix2"], ax[0]["xixx"]
        self.ety2Staggr, self.etyyStaggr = ax[1]["xix2"], ax[1]["xixx"]
        self.ztz2Staggr, self.ztzzStaggr = ax[2]["xix2"], ax[2]["xixx"]
        self.dXi, self.dEt, self.dZt = (1.0 / (k - 1) for k in self.n)

    def shape(self):
        return tuple(k + 2 for k in self.n)


def _dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Multigrid:
    """multigrid_d3 on the device: ``mgSolve`` plus the protected operators for parity tests."""

    def __init__(self, grid: Grid, p: Parser, walls=None):
        self.lib = _L.load()
        self.grid, self.p = grid, p
        self._keep = []
        d = _L.SrMgDesc()
        d.xInd, d.yInd, d.zInd = p.xInd, p.yInd, p.zInd
        d.vcDepth, d.vcCount = p.vcDepth, p.vcCount
        d.preSmooth, d.postSmooth = p.preSmooth, p.postSmooth
        d.gsSmooth, d.resType = int(p.gsSmooth), p.resType
        d.xPer, d.yPer, d.zPer = int(p.xPer), int(p.yPer), int(p.zPer)
        d.testPoisson, d.semantics, d.precision = int(p.testPoisson), p.semantics, p.precision
        d.padMode, d.coarseProject, d.coarseDirect = p.padMode, p.coarseProject, p.coarseDirect
        d.innerPreSmooth, d.innerPostSmooth = p.innerPreSmooth, p.innerPostSmooth
        d.mgTolerance, d.relTolerance = p.mgTolerance, p.relTolerance
        d.dXi, d.dEt, d.dZt = grid.dXi, grid.dEt, grid.dZt
        for name, arr in (("xix2", grid.xix2Staggr), ("xixx", grid.xixxStaggr), ("ety2", grid.ety2Staggr),
                          ("etyy", grid.etyyStaggr), ("ztz2", grid.ztz2Staggr), ("ztzz", grid.ztzzStaggr)):
            a = np.ascontiguousarray(arr, dtype=np.float64)
            self._keep.append(a)
            setattr(d, name, _dptr(a))
        for q, w in enumerate(walls or [None] * 6):
            if w is not None:
                a = np.ascontiguousarray(w, dtype=np.float64)
                self._keep.append(a)
                d.wall[q] = _dptr(a)
        h = C.c_void_p()
        _L.check(self.lib.sr_mg_create(C.byref(d), C.byref(h)))
        self.h = h
        self.desc = d

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.sr_mg_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- poisson::mgSolve -------------------------------------------------------------------
    def mgSolve(self, inFn, rhs):
        """In place on ``inFn`` (numpy, padded) like poisson::mgSolve(plainsf&, const plainsf&).
        Returns the stats (cycles, converged, residual, solve_ms)."""
        assert inFn.shape == self.grid.shape() and rhs.shape == self.grid.shape()
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        assert inFn.dtype == np.float64 and inFn.flags["C_CONTIGUOUS"]
        st = _L.SrMgStats()
        code = self.lib.sr_mg_solve(self.h, inFn.ctypes.data, rhs.ctypes.data, 0, C.byref(st))
        if code not in (_L.SR_OK, _L.SR_ERR_NOT_CONVERGED):
            _L.check(code)
        return st

    def mgSolve_device(self, p_ptr, rhs_ptr):
        """Solve on device buffers (HBM-resident padded arrays, e.g. torch tensors' data_ptr())."""
        st = _L.SrMgStats()
        code = self.lib.sr_mg_solve(self.h, C.c_void_p(p_ptr), C.c_void_p(rhs_ptr), 1, C.byref(st))
        if code not in (_L.SR_OK, _L.SR_ERR_NOT_CONVERGED):
            _L.check(code)
        return st

    def residual_device(self, p_ptr, rhs_ptr, r_ptr=None):
        """Relative residual (the stop rule's norm) of a device guess; optionally f - L p."""
        rel = C.c_double()
        _L.check(self.lib.sr_mg_residual(self.h, C.c_void_p(p_ptr), C.c_void_p(rhs_ptr),
                                         C.c_void_p(r_ptr) if r_ptr else None, C.byref(rel)))
        return rel.value

    def history(self, cap=1024):
        out = np.zeros(cap)
        n = self.lib.sr_mg_history(self.h, _dptr(out), cap)
        return out[:max(n, 0)].copy()

    # ---- protected operators (multigrid_d3) ---------------------------------------------------
    @property
    def vLevel(self):
        v = C.c_int()
        _L.check(self.lib.sr_mg_get_level(self.h, C.byref(v)))
        return v.value

    @vLevel.setter
    def vLevel(self, level):
        _L.check(self.lib.sr_mg_set_level(self.h, int(level)))

    def set_zeroBC(self, z):
        _L.check(self.lib.sr_mg_set_zero_bc(self.h, int(bool(z))))

    def dims(self, level):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _L.check(self.lib.sr_mg_level_dims(self.h, int(level), C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def upload(self, which, level, arr):
        nx, ny, nz = self.dims(level)
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == (nx + 2, ny + 2, nz + 2)
        _L.check(self.lib.sr_mg_upload(self.h, int(which), int(level), _dptr(a)))

    def download(self, which, level):
        nx, ny, nz = self.dims(level)
        a = np.zeros((nx + 2, ny + 2, nz + 2))
        _L.check(self.lib.sr_mg_download(self.h, int(which), int(level), _dptr(a)))
        return a

    def smooth(self, count):
        _L.check(self.lib.sr_mg_op_smooth(self.h, int(count)))

    def computeResidual(self):
        _L.check(self.lib.sr_mg_op_compute_residual(self.h))

    def coarsen(self):
        _L.check(self.lib.sr_mg_op_coarsen(self.h))

    def prolong(self):
        _L.check(self.lib.sr_mg_op_prolong(self.h))

    def solve(self):
        _L.check(self.lib.sr_mg_op_solve_coarsest(self.h))

    def imposeBC(self):
        _L.check(self.lib.sr_mg_op_impose_bc(self.h))

    def updatePads(self, which):
        _L.check(self.lib.sr_mg_op_update_pads(self.h, int(which)))

    def computeError(self, normOrder):
        out = C.c_double()
        _L.check(self.lib.sr_mg_op_compute_error(self.h, int(normOrder), C.byref(out)))
        return out.value

    def vCycle(self):
        _L.check(self.lib.sr_mg_op_vcycle(self.h))

    # ---- measurement --------------------------------------------------------------------------
    def time_kernel(self, kernel, level, reps):
        out = C.c_double()
        _L.check(self.lib.sr_mg_time_kernel(self.h, int(kernel), int(level), int(reps), C.byref(out)))
        return out.value

    def info(self, what):
        v = C.c_long()
        _L.check(self.lib.sr_mg_info(self.h, int(what), C.byref(v)))
        return v.value

    def kernel_bytes(self, kernel, level):
        out = C.c_double()
        _L.check(self.lib.sr_mg_kernel_bytes(self.h, int(kernel), int(level), C.byref(out)))
        return out.value

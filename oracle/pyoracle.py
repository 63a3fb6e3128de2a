"""ctypes view of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py import this
module, and only as the checker or the CPU baseline.  The product package ``saras_amd`` never
imports it.

Arrays follow the reference layout: padded (Nx+2, Ny+2, Nz+2) float64, C order, k fastest,
element [0] along an axis is logical index -1.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")

SMOOTH_LEX, SMOOTH_RB, SMOOTH_JACOBI, SMOOTH_JACOBI_RB = 0, 1, 2, 3
SEM_REFERENCE, SEM_PROJECTED = 0, 1
P, R, T, S = 0, 1, 2, 3  # pressureData, residualData, tmpDataArray, smoothedPres


class Params(C.Structure):
    _fields_ = [
        ("xInd", C.c_int), ("yInd", C.c_int), ("zInd", C.c_int),
        ("vcDepth", C.c_int), ("vcCount", C.c_int),
        ("preSmooth", C.c_int), ("postSmooth", C.c_int),
        ("resType", C.c_int), ("smoother", C.c_int),
        ("xPer", C.c_int), ("yPer", C.c_int), ("zPer", C.c_int),
        ("testPoisson", C.c_int), ("semantics", C.c_int),
        ("mgTolerance", C.c_double), ("relTolerance", C.c_double),
        ("dXi", C.c_double), ("dEt", C.c_double), ("dZt", C.c_double),
        ("padMode", C.c_int), ("coarseProject", C.c_int),
    ]


def build():
    """Compile the oracle libraries (gcc, a few seconds)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_libs = {}


def _bind(L):
    """argtypes of the multigrid entry points (oracle.h)."""
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    L.or_grid_staggr.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, dp, dp, dp, dp]
    L.or_grid_colloc.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, dp, dp, dp, dp]
    L.or_mg_create.restype = vp
    L.or_mg_create.argtypes = [C.POINTER(Params)] + [dp] * 12
    L.or_mg_destroy.argtypes = [vp]
    L.or_mg_array.restype = dp
    L.or_mg_array.argtypes = [vp, C.c_int, C.c_int]
    L.or_mg_dims.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    for f in ("or_mg_set_level", "or_mg_set_zero_bc", "or_mg_smooth", "or_mg_update_pads"):
        getattr(L, f).argtypes = [vp, C.c_int]
    L.or_mg_get_level.restype = C.c_int
    L.or_mg_get_level.argtypes = [vp]
    for f in ("or_mg_compute_residual", "or_mg_solve_coarsest", "or_mg_coarsen", "or_mg_prolong",
              "or_mg_impose_bc", "or_mg_vcycle"):
        getattr(L, f).argtypes = [vp]
    L.or_mg_compute_error.restype = C.c_double
    L.or_mg_compute_error.argtypes = [vp, C.c_int]
    L.or_mg_solve.restype = C.c_int
    L.or_mg_solve.argtypes = [vp, dp, dp, dp, C.c_int]
    L.or_null_vector_1d.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp, dp]
    L.or_mg_project_rhs.argtypes = [vp, dp]
    L.or_mg_metric.restype = dp
    L.or_mg_metric.argtypes = [vp, C.c_int, C.c_int]
    L.or_mg_coef.restype = C.c_double
    L.or_mg_coef.argtypes = [vp, C.c_int, C.c_int]
    L.or_mg_max_count.restype = C.c_int
    L.or_mg_max_count.argtypes = [vp]
    return L


def lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(_BUILD, name)
    srcs = [os.path.join(os.path.dirname(os.path.abspath(__file__)), f)
            for f in ("mg_oracle.c", "field_oracle.c", "oracle.h")]
    if not os.path.exists(path) or any(os.path.getmtime(f) > os.path.getmtime(path) for f in srcs if os.path.exists(f)):
        build()
    _libs[name] = _bind(C.CDLL(path))
    return _libs[name]


def load(path):
    """A separately built copy of the oracle (e.g. a deliberately mutated one in a test), bound
    like lib(); pass it as `library=` to OracleMG / OracleTS."""
    return _bind(C.CDLL(os.path.abspath(path)))


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def grid_axis(n, length=1.0, grid_type=1, beta=1.0, kind="staggr", fast=False):
    """Node (staggr, padded n+2) or face (colloc, padded n+1) coordinates and metrics."""
    L = lib(fast)
    m = n + 2 if kind == "staggr" else n + 1
    x, xi_x, xixx, xix2 = (np.zeros(m) for _ in range(4))
    fn = L.or_grid_staggr if kind == "staggr" else L.or_grid_colloc
    fn(n, float(length), int(grid_type), float(beta), _ptr(x), _ptr(xi_x), _ptr(xixx), _ptr(xix2))
    return {"x": x, "xi_x": xi_x, "xixx": xixx, "xix2": xix2}


class OracleMG:
    """The restated multigrid_d3 object: level hierarchies plus its protected operators."""

    def __init__(self, params, metrics, walls=None, fast=False, library=None):
        self.L = library if library is not None else lib(fast)
        self.params = params
        self._keep = []
        ms = []
        for key in ("xix2", "xixx", "ety2", "etyy", "ztz2", "ztzz"):
            a = np.ascontiguousarray(metrics[key], dtype=np.float64)
            self._keep.append(a)
            ms.append(_ptr(a))
        ws = []
        for w in (walls or [None] * 6):
            if w is None:
                ws.append(None)
            else:
                a = np.ascontiguousarray(w, dtype=np.float64)
                self._keep.append(a)
                ws.append(_ptr(a))
        self.h = self.L.or_mg_create(C.byref(params), *ms, *ws)
        if not self.h:
            raise RuntimeError("or_mg_create failed")

    def close(self):
        if self.h:
            self.L.or_mg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dims(self, level):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self.L.or_mg_dims(self.h, level, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def array(self, which, level):
        """numpy view (no copy) of a level array; valid until the next Jacobi swap."""
        nx, ny, nz = self.dims(level)
        p = self.L.or_mg_array(self.h, which, level)
        return np.ctypeslib.as_array(p, shape=(nx + 2, ny + 2, nz + 2))

    def metric(self, which, level):
        n = self.dims(level)[which // 2]
        return np.ctypeslib.as_array(self.L.or_mg_metric(self.h, which, level), shape=(n + 2,)).copy()

    def coef(self, which, level):
        return self.L.or_mg_coef(self.h, which, level)

    @property
    def level(self):
        return self.L.or_mg_get_level(self.h)

    @level.setter
    def level(self, v):
        self.L.or_mg_set_level(self.h, int(v))

    def set_zero_bc(self, z):
        self.L.or_mg_set_zero_bc(self.h, int(z))

    def compute_residual(self):
        self.L.or_mg_compute_residual(self.h)

    def smooth(self, n):
        self.L.or_mg_smooth(self.h, int(n))

    def solve_coarsest(self):
        self.L.or_mg_solve_coarsest(self.h)

    def coarsen(self):
        self.L.or_mg_coarsen(self.h)

    def prolong(self):
        self.L.or_mg_prolong(self.h)

    def compute_error(self, norm):
        return self.L.or_mg_compute_error(self.h, int(norm))

    def impose_bc(self):
        self.L.or_mg_impose_bc(self.h)

    def update_pads(self, which):
        self.L.or_mg_update_pads(self.h, int(which))

    def vcycle(self):
        self.L.or_mg_vcycle(self.h)

    def max_count(self):
        return self.L.or_mg_max_count(self.h)

    def project_rhs(self, rhs):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        self.L.or_mg_project_rhs(self.h, _ptr(rhs))
        return rhs

    def solve(self, p, rhs, max_hist=256):
        p = np.ascontiguousarray(p, dtype=np.float64).copy()
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        hist = np.zeros(max_hist)
        n = self.L.or_mg_solve(self.h, _ptr(p), _ptr(rhs), _ptr(hist), max_hist)
        return p, hist[:min(n, max_hist)].copy()


def null_vector_1d(n, periodic, h, d2, d1):
    L = lib()
    w = np.zeros(n)
    d2 = np.ascontiguousarray(d2, dtype=np.float64)
    d1 = np.ascontiguousarray(d1, dtype=np.float64)
    L.or_null_vector_1d(n, int(periodic), float(h), _ptr(d2), _ptr(d1), _ptr(w))
    return w


# ------------------------------------------------------------------------------------------------
# hydro time step (field_oracle.c)
# ------------------------------------------------------------------------------------------------
class TsParams(C.Structure):
    _fields_ = [
        ("xInd", C.c_int), ("yInd", C.c_int), ("zInd", C.c_int),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("xGrid", C.c_int), ("yGrid", C.c_int), ("zGrid", C.c_int),
        ("betaX", C.c_double), ("betaY", C.c_double), ("betaZ", C.c_double),
        ("xPer", C.c_int), ("yPer", C.c_int), ("zPer", C.c_int),
        ("probType", C.c_int),
        ("Re", C.c_double), ("dt", C.c_double), ("cnTolerance", C.c_double),
        ("vcDepth", C.c_int), ("vcCount", C.c_int), ("preSmooth", C.c_int), ("postSmooth", C.c_int),
        ("resType", C.c_int), ("smoother", C.c_int),
        ("mgTolerance", C.c_double),
        ("scalar", C.c_int), ("rbcType", C.c_int), ("Ra", C.c_double), ("Pr", C.c_double),
        ("semantics", C.c_int), ("relTolerance", C.c_double),
    ]


VX, VY, VZ, PRES, PCORR, MGRHS, HX, HY, HZ, GX, GY, GZ, DIV, TEMP, TRHS = range(15)


def _c(a):
    """float64 C-contiguous view or copy of a read-only input"""
    return np.ascontiguousarray(a, dtype=np.float64)


def _ts_bind(L):
    if getattr(L, "_ts_bound", False):
        return
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    L.or_ts_create.restype = vp
    L.or_ts_create.argtypes = [C.POINTER(TsParams)]
    L.or_ts_destroy.argtypes = [vp]
    L.or_ts_array.restype = dp
    L.or_ts_array.argtypes = [vp, C.c_int]
    L.or_ts_shape.argtypes = [vp, C.c_int, ip, ip, ip]
    L.or_ts_metric.restype = dp
    L.or_ts_metric.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.or_ts_deriv.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
    for f in ("or_ts_compute_diff", "or_ts_compute_nlin", "or_ts_divergence", "or_ts_init_tgv",
              "or_ts_build_rhs"):
        getattr(L, f).argtypes = [vp]
    L.or_ts_gradient.argtypes = [vp, C.c_int]
    L.or_ts_impose_vbc.argtypes = [vp, C.c_int]
    L.or_ts_energy.restype = C.c_double
    L.or_ts_energy.argtypes = [vp]
    L.or_ts_max_div.restype = C.c_double
    L.or_ts_max_div.argtypes = [vp]
    L.or_ts_solve_v.restype = C.c_int
    L.or_ts_solve_v.argtypes = [vp, C.c_int]
    L.or_ts_time_advance.restype = C.c_int
    L.or_ts_time_advance.argtypes = [vp]
    L.or_ts_jacobi_iters.restype = C.c_int
    L.or_ts_jacobi_iters.argtypes = [vp, C.c_int]
    L.or_ts_mg_cycles.restype = C.c_int
    L.or_ts_mg_cycles.argtypes = [vp]
    L.or_ts_init_rbc.argtypes = [vp, C.c_ulonglong, C.c_double]
    L.or_ts_impose_tbc.argtypes = [vp]
    L.or_ts_build_trhs.argtypes = [vp]
    L.or_ts_solve_t.restype = C.c_int
    L.or_ts_solve_t.argtypes = [vp]
    L.or_ts_nusselt.restype = C.c_double
    L.or_ts_nusselt.argtypes = [vp]
    L.or_fd_deriv.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp]
    L.or_fd_diff.argtypes = [vp, C.c_int, dp, dp]
    L.or_fd_nlin.argtypes = [vp, C.c_int, dp, C.POINTER(dp), dp]
    L.or_fd_gradient.argtypes = [vp, dp, dp, dp, dp]
    L.or_fd_divergence.argtypes = [vp, dp, dp, dp, dp]
    L.or_fd_bc.argtypes = [vp, C.c_int, ip, dp, dp]
    L._ts_bound = True


def ts_params(ind=(5, 5, 5), L=(2 * np.pi,) * 3, domain="PPP", mesh="UUU", beta=(1.0, 1.0, 1.0), Re=1600.0,
              dt=1e-3, cn_tol=1e-6, depth=None, vc_count=5, pre=4, post=4, res_type=1, smoother=SMOOTH_RB,
              mg_tol=1e-6, prob_type=2, scalar=False, rbc_type=2, Ra=1e7, Pr=1.0, semantics=SEM_REFERENCE,
              rel_tol=1e-10):
    p = TsParams()
    p.scalar, p.rbcType, p.Ra, p.Pr = int(scalar), rbc_type, Ra, Pr
    p.semantics, p.relTolerance = semantics, rel_tol
    p.xInd, p.yInd, p.zInd = ind
    p.Lx, p.Ly, p.Lz = L
    p.xGrid, p.yGrid, p.zGrid = (1 if c == "U" else 2 for c in mesh)
    p.betaX, p.betaY, p.betaZ = beta
    p.xPer, p.yPer, p.zPer = (int(c == "P") for c in domain)
    p.probType = prob_type
    p.Re, p.dt, p.cnTolerance = Re, dt, cn_tol
    p.vcDepth = depth if depth is not None else min(ind) - 1
    p.vcCount, p.preSmooth, p.postSmooth, p.resType, p.smoother = vc_count, pre, post, res_type, smoother
    p.mgTolerance = mg_tol
    return p


class OracleTS:
    """The restated hydro_d3 time step (fields, stencils, Jacobi viscous solve, projection)."""

    def __init__(self, params, fast=False, library=None):
        self.L = library if library is not None else lib(fast)
        _ts_bind(self.L)
        self.params = params
        self.h = self.L.or_ts_create(C.byref(params))

    def close(self):
        if self.h:
            self.L.or_ts_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, which):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self.L.or_ts_shape(self.h, which, C.byref(a), C.byref(b), C.byref(c))
        return a.value + 2, b.value + 2, c.value + 2

    def array(self, which):
        return np.ctypeslib.as_array(self.L.or_ts_array(self.h, which), shape=self.shape(which))

    def metric(self, dim, node, quantity):
        n = (1 << (self.params.xInd, self.params.yInd, self.params.zInd)[dim]) + 1
        m = n + 2 if node else n + 1
        return np.ctypeslib.as_array(self.L.or_ts_metric(self.h, dim, int(node), quantity), shape=(m,)).copy()

    def deriv(self, which, order, dim):
        out = np.zeros(self.shape(which))
        self.L.or_ts_deriv(self.h, which, order, dim, _ptr(out))
        return out

    def init_tgv(self):
        self.L.or_ts_init_tgv(self.h)

    def compute_diff(self):
        self.L.or_ts_compute_diff(self.h)

    def compute_nlin(self):
        self.L.or_ts_compute_nlin(self.h)

    def gradient(self, from_pp=False):
        self.L.or_ts_gradient(self.h, int(from_pp))

    def divergence(self):
        self.L.or_ts_divergence(self.h)

    def impose_vbc(self, c):
        self.L.or_ts_impose_vbc(self.h, c)

    def build_rhs(self):
        self.L.or_ts_build_rhs(self.h)

    def solve_v(self, c):
        return self.L.or_ts_solve_v(self.h, c)

    def time_advance(self):
        r = self.L.or_ts_time_advance(self.h)
        if r != 0:
            raise RuntimeError("Jacobi iterations did not converge")

    def energy(self):
        return self.L.or_ts_energy(self.h)

    def max_div(self):
        return self.L.or_ts_max_div(self.h)

    def jacobi_iters(self, c):
        return self.L.or_ts_jacobi_iters(self.h, c)

    def mg_cycles(self):
        return self.L.or_ts_mg_cycles(self.h)

    def init_rbc(self, seed=0, amp=1e-3):
        self.L.or_ts_init_rbc(self.h, seed, amp)

    def impose_tbc(self):
        self.L.or_ts_impose_tbc(self.h)

    def build_trhs(self):
        self.L.or_ts_build_trhs(self.h)

    def solve_t(self):
        return self.L.or_ts_solve_t(self.h)

    def nusselt(self):
        return self.L.or_ts_nusselt(self.h)

    # field operators on caller arrays (loc 0 node, 1..3 the faces of Vx..Vz); arrays are float64,
    # C-contiguous, of the location's padded shape, and are modified in place where written
    def loc_shape(self, loc):
        return self.shape(PRES if loc == 0 else loc - 1)

    def fd_deriv(self, loc, order, dim, F):
        out = np.zeros(self.loc_shape(loc))
        self.L.or_fd_deriv(self.h, loc, order, dim, _ptr(_c(F)), _ptr(out))
        return out

    def fd_diff(self, loc, F, H):
        self.L.or_fd_diff(self.h, loc, _ptr(_c(F)), _ptr(H))

    def fd_nlin(self, loc, F, V, H):
        Vc = [_c(v) for v in V]
        arr = (C.POINTER(C.c_double) * 3)(*[_ptr(v) for v in Vc])
        self.L.or_fd_nlin(self.h, loc, _ptr(_c(F)), arr, _ptr(H))

    def fd_gradient(self, F, G):
        self.L.or_fd_gradient(self.h, _ptr(_c(F)), *[_ptr(g) for g in G])

    def fd_divergence(self, V):
        div = np.zeros(self.shape(PRES))
        Vc = [_c(v) for v in V]
        self.L.or_fd_divergence(self.h, *[_ptr(v) for v in Vc], _ptr(div))
        return div

    def fd_bc(self, loc, kinds, values, F):
        k = (C.c_int * 6)(*kinds)
        v = np.ascontiguousarray(values, dtype=np.float64)
        self.L.or_fd_bc(self.h, loc, k, _ptr(v), _ptr(F))

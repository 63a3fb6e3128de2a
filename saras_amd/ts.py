"""Python mirror of hydro_d3 (src/solvers/hydro/hydro_d3.cc) over the C ABI of include/saras_ts.h.

Every field stays in HBM for the life of the object; `download` copies one padded field (the
reference layout: C order, k fastest, index -1 first) to the host.  Method names follow the
reference: timeAdvance, solveVx/Vy/Vz, imposeUBCs/VBCs/WBCs, and the tseries diagnostics."""
import ctypes as C

import numpy as np

from . import _lib as _L

VX, VY, VZ, P, PP, MGRHS = _L.SR_TS_VX, _L.SR_TS_VY, _L.SR_TS_VZ, _L.SR_TS_P, _L.SR_TS_PP, _L.SR_TS_MGRHS
HX, HY, HZ, DIV = _L.SR_TS_HX, _L.SR_TS_HY, _L.SR_TS_HZ, _L.SR_TS_DIV
T, TRHS = _L.SR_TS_T, _L.SR_TS_TRHS
PP1, PP2 = _L.SR_TS_PP1, _L.SR_TS_PP2


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def ts_desc(ind=(7, 7, 7), L=(2 * np.pi,) * 3, domain="PPP", mesh="UUU", beta=(1.0, 1.0, 1.0), Re=1600.0,
            dt=1e-3, cn_tol=1e-6, depth=None, vc_count=5, pre=4, post=4, res_type=1, mg_tol=1e-6, prob_type=2,
            scalar=False, rbc_type=2, Ra=1e7, Pr=1.0, semantics=_L.SR_SEM_REFERENCE, rel_tol=1e-10,
            precision=_L.SR_PREC_FP64, coarse_direct=0, gs_smooth=1, inner=(0, 0), pad_mode=0):
    """The parser fields of a hydro (or, with scalar=True, scalar) run (input/parameters.yaml keys;
    defaults: the Taylor-Green case)."""
    d = _L.SrTsDesc()
    d.scalar, d.rbcType, d.Ra, d.Pr = int(scalar), rbc_type, Ra, Pr
    d.semantics, d.relTolerance, d.precision = semantics, rel_tol, precision
    d.coarseDirect = int(coarse_direct)
    d.innerPreSmooth, d.innerPostSmooth = (int(k) for k in inner)
    d.padMode = int(pad_mode)
    d.xInd, d.yInd, d.zInd = ind
    d.Lx, d.Ly, d.Lz = L
    d.xGrid, d.yGrid, d.zGrid = (1 if c == "U" else 2 for c in mesh)
    d.betaX, d.betaY, d.betaZ = beta
    d.xPer, d.yPer, d.zPer = (int(c == "P") for c in domain)
    d.probType = prob_type
    d.Re, d.dt, d.cnTolerance = Re, dt, cn_tol
    d.vcDepth = depth if depth is not None else min(ind) - 1
    d.vcCount, d.preSmooth, d.postSmooth, d.resType, d.gsSmooth = vc_count, pre, post, res_type, int(gs_smooth)
    d.mgTolerance = mg_tol
    return d


class Hydro:
    """hydro_d3 on the device: construct, set an initial condition, then timeAdvance()."""

    def __init__(self, desc):
        self.L = _L.load()
        self.desc = desc
        h = C.c_void_p()
        _L.check(self.L.sr_ts_create(C.byref(desc), C.byref(h)))
        self.h = h
        self.stats = _L.SrTsStats()

    def close(self):
        if self.h:
            self.L.sr_ts_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- fields -----------------------------------------------------------------------------
    def kernel_bytes(self, kernel):
        b = C.c_double()
        _L.check(self.L.sr_ts_kernel_bytes(self.h, kernel, C.byref(b)))
        return b.value

    def time_kernel(self, kernel, reps=10):
        ms = C.c_double()
        _L.check(self.L.sr_ts_time_kernel(self.h, kernel, reps, C.byref(ms)))
        return ms.value

    def device_ptr(self, which):
        """Device address of a field's padded array (valid while the handle lives)."""
        p = C.c_void_p()
        _L.check(self.L.sr_ts_field_ptr(self.h, which, C.byref(p)))
        return p.value

    def shape(self, which):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _L.check(self.L.sr_ts_field_dims(self.h, which, C.byref(a), C.byref(b), C.byref(c)))
        return a.value + 2, b.value + 2, c.value + 2

    def download(self, which):
        out = np.empty(self.shape(which))
        _L.check(self.L.sr_ts_download(self.h, which, _ptr(out)))
        return out

    def upload(self, which, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.shape(which):
            raise ValueError(f"field {which}: shape {a.shape} != {self.shape(which)}")
        _L.check(self.L.sr_ts_upload(self.h, which, _ptr(a)))

    # -- the step and its pieces ------------------------------------------------------------
    def initTGV(self):
        _L.check(self.L.sr_ts_init_tgv(self.h))

    def timeAdvance(self):
        """One step; returns this step's statistics (a new record every call)."""
        st = _L.SrTsStats()
        _L.check(self.L.sr_ts_time_advance(self.h, C.byref(st)))
        self.stats = st
        return st

    def deriv(self, which, order, dim):
        out = np.empty(self.shape(which))
        _L.check(self.L.sr_ts_deriv(self.h, which, order, dim, _ptr(out)))
        return out

    def buildRHS(self):
        _L.check(self.L.sr_ts_build_rhs(self.h))

    def solveV(self, component):
        it = C.c_int()
        _L.check(self.L.sr_ts_solve_v(self.h, component, C.byref(it)))
        return it.value

    def solveVx(self):
        return self.solveV(0)

    def solveVy(self):
        return self.solveV(1)

    def solveVz(self):
        return self.solveV(2)

    def imposeVBC(self, component):
        _L.check(self.L.sr_ts_impose_vbc(self.h, component))

    def divergence(self):
        _L.check(self.L.sr_ts_divergence(self.h))
        return self.download(DIV)

    def energy(self):
        e = C.c_double()
        _L.check(self.L.sr_ts_energy(self.h, C.byref(e)))
        return e.value

    def maxDivergence(self):
        m = C.c_double()
        _L.check(self.L.sr_ts_max_div(self.h, C.byref(m)))
        return m.value

    # -- scalar_d3 ----------------------------------------------------------------------------
    def initRBC(self, seed=0, amp=1e-3):
        _L.check(self.L.sr_ts_init_rbc(self.h, seed, amp))

    def imposeTBCs(self):
        _L.check(self.L.sr_ts_impose_tbc(self.h))

    def buildTRHS(self):
        _L.check(self.L.sr_ts_build_trhs(self.h))

    def solveT(self):
        it = C.c_int()
        _L.check(self.L.sr_ts_solve_t(self.h, C.byref(it)))
        return it.value

    def nusselt(self):
        v = C.c_double()
        _L.check(self.L.sr_ts_nusselt(self.h, C.byref(v)))
        return v.value

    def coefficients(self):
        """(nu, kappa) of the run: nu = 1/Re for hydro, by rbcType for scalar runs (kappa 0 for hydro)."""
        nu, kappa = C.c_double(), C.c_double()
        _L.check(self.L.sr_ts_coefficients(self.h, C.byref(nu), C.byref(kappa)))
        return nu.value, kappa.value

    # -- learned initial guess of the pressure solve -------------------------------------------
    def attach_guess(self, net, train=True):
        """net: saras_amd.guess.GuessNet whose solver grid is this run's node grid (or None)."""
        _L.check(self.L.sr_ts_attach_guess(self.h, net.h if net is not None else None, int(train)))
        self._guess = net

    def solve_history(self, cap=256):
        """Relative residuals after each cycle of the last pressure solve (sr_mg_history)."""
        m = C.c_void_p()
        _L.check(self.L.sr_ts_pressure_solver(self.h, C.byref(m)))
        out = np.zeros(cap)
        n = self.L.sr_mg_history(m, out.ctypes.data_as(C.POINTER(C.c_double)), cap)
        return out[:max(n, 0)].copy()

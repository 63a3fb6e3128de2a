"""The benchmark configurations at their full sizes on the device.

* C4: the 257^3 stretched-grid Poisson case (PPN, tanh z beta = 1, V(2,2), 1e-10 relative
  residual) over several seeds of the 8-mode Fourier right-hand side: plain FP64, FP32 cycles under
  FP64 refinement and the GCR-accelerated mixed solve; the mixed solutions' max-abs and L2 errors
  against the FP64 solve, and at seed 0 all three against the oracle's FP64 multigrid (red-black
  order, the device's).
* C3: the Taylor-Green vortex on 129^3 nodes (128 unique periodic cells, L = 2 pi, Re = 1600,
  dt = 1e-3): whole steps against the oracle's step, fields and the kinetic-energy trace.

Bounds: the FP64 solve matches the oracle's to 1e-12 relative (same cycle order); refinement follows
the FP64 iterates (L2 difference <= 1e-10 relative); GCR stops at another point inside the same
1e-10 residual ball (L2 difference <= 1e-8 relative, bounded by the residual times the operator's
conditioning); the time-step fields match at 1e-12 like the 33^3 cases, the energy to the rounding
bound of its sequential sum."""
import numpy as np
import pytest

from oracle import pyoracle as po
from helpers import rel_close

pytestmark = pytest.mark.gpu


def _centred(a):
    c = a[1:-1, 1:-1, 1:-1]
    return c - c.mean()


def _diffs(a, b):
    d = _centred(a) - _centred(b)
    ref = _centred(b)
    return float(np.max(np.abs(d)) / np.max(np.abs(ref))), float(np.linalg.norm(d) / np.linalg.norm(ref))


def test_c4_mixed_precision_257_cubed_seeds():
    from saras_amd import mg, workloads
    from saras_amd._lib import DeviceArray
    p64, grid, f0 = workloads.stretched_poisson(ind=8, beta=1.0, seed=0)
    solvers = {}
    for name, prec, direct in (("fp64", mg._L.SR_PREC_FP64, 0), ("refinement", mg._L.SR_PREC_MIXED, 0),
                               ("gcr", mg._L.SR_PREC_MIXED_GCR, 1)):
        p = workloads.stretched_poisson(ind=8, beta=1.0, seed=0)[0]
        p.precision, p.coarseDirect = prec, direct
        solvers[name] = mg.Multigrid(grid, p)
    d_rhs, d_p = DeviceArray(f0.shape), DeviceArray(f0.shape)
    table = []
    sols0 = {}
    for seed in (0, 3, 7, 12):
        f = f0 if seed == 0 else workloads.fourier_rhs(grid.xStaggr, grid.yStaggr, grid.zStaggr, seed)
        d_rhs.upload(f)
        sols = {}
        for name, s in solvers.items():
            d_p.upload(np.zeros(f.shape))
            st = s.mgSolve_device(d_p.ptr, d_rhs.ptr)
            assert st.converged and st.residual < 1e-10, (seed, name, st.residual)
            sols[name] = d_p.download()
        ref_abs, ref_l2 = _diffs(sols["refinement"], sols["fp64"])
        gcr_abs, gcr_l2 = _diffs(sols["gcr"], sols["fp64"])
        table.append((seed, ref_abs, ref_l2, gcr_abs, gcr_l2))
        assert ref_l2 <= 1e-10 and ref_abs <= 1e-10, (seed, ref_abs, ref_l2)
        assert gcr_l2 <= 1e-8 and gcr_abs <= 1e-8, (seed, gcr_abs, gcr_l2)
        if seed == 0:
            sols0 = sols
    print("seed, refinement max-abs / L2 rel. diff, GCR max-abs / L2 rel. diff vs FP64:")
    for row in table:
        print("  %2d  %.2e %.2e  %.2e %.2e" % row)
    for s in solvers.values():
        s.close()
    # seed 0 against the oracle's FP64 multigrid from zero
    o = _oracle_solver(p64, grid)
    po_sol, hist = o.solve(np.zeros(f0.shape), f0)
    o.close()
    assert hist[-1] < 1e-10
    a64, l64 = _diffs(sols0["fp64"], po_sol)
    assert a64 <= 1e-12, a64
    for name, bound in (("refinement", 1e-10), ("gcr", 1e-8)):
        a, l2 = _diffs(sols0[name], po_sol)
        assert a <= bound and l2 <= bound, (name, a, l2)


def _oracle_solver(p, g):
    prm = po.Params()
    prm.xInd, prm.yInd, prm.zInd = p.xInd, p.yInd, p.zInd
    prm.vcDepth, prm.vcCount = p.vcDepth, 60
    prm.preSmooth, prm.postSmooth, prm.resType = p.preSmooth, p.postSmooth, 2
    prm.smoother = po.SMOOTH_RB
    prm.xPer, prm.yPer, prm.zPer = int(p.xPer), int(p.yPer), int(p.zPer)
    prm.testPoisson, prm.semantics = 0, po.SEM_PROJECTED
    prm.mgTolerance, prm.relTolerance = p.mgTolerance, p.relTolerance
    prm.dXi, prm.dEt, prm.dZt = g.dXi, g.dEt, g.dZt
    met = {"xix2": g.xix2Staggr, "xixx": g.xixxStaggr, "ety2": g.ety2Staggr, "etyy": g.etyyStaggr,
           "ztz2": g.ztz2Staggr, "ztzz": g.ztzzStaggr}
    return po.OracleMG(prm, met, fast=False)


def test_c3_taylor_green_129_cubed_steps():
    from saras_amd import ts
    kw = dict(ind=(7, 7, 7), L=(2 * np.pi,) * 3, domain="PPP", mesh="UUU", Re=1600.0, dt=1e-3)
    o = po.OracleTS(po.ts_params(**kw))
    g = ts.Hydro(ts.ts_desc(**kw))
    o.init_tgv()
    g.initTGV()
    e0 = o.energy()
    for step in range(3):
        o.time_advance()
        st = g.timeAdvance()
        assert [st.jacobi_iters[c] for c in range(3)] == [o.jacobi_iters(c) for c in range(3)], step
        assert st.mg_cycles == o.mg_cycles(), step
        vs = max(float(np.max(np.abs(o.array(c)))) for c in range(3))
        for c in range(3):
            ok, err, s = rel_close(g.download(c), o.array(c), 1e-12, scale=vs)
            assert ok, (step, c, err)
        # the energy is a sum over 129^3 = 2.1e6 nodes: the reference (and the oracle) add them one
        # after the other (tseries.cc:158-169), whose rounding is bounded by n eps |E| = 2.4e-10 |E|;
        # the device adds them as a tree.  The bar is that bound (measured 4e-11 at the first step).
        e_o, e_g = o.energy(), g.energy()
        assert abs(e_o - e_g) <= 129 ** 3 * 2.2e-16 * abs(e_o), (step, e_o, e_g)
    # the energy decays at the viscous rate of the vortex, E ~ E0 exp(-6 t / Re) at early times
    assert abs(e_g / e0 - np.exp(-6 * 3e-3 / 1600.0)) < 1e-4
    g.close()
    o.close()

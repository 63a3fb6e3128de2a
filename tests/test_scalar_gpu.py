"""Parity of the gfx950 scalar_d3 step (temperature, buoyancy, RBC walls) with the CPU oracle.

Bar as for the hydro step: the initial state and the temperature walls bit for bit; buoyant
nseRHS, tmpRHS and the Jacobi temperature solve within 1e-12 relative; whole steps (fields, Jacobi
and V-cycle counts, energy and Nusselt number) within 1e-12, pressures compared with their means
removed (the pressure is fixed up to a constant)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from helpers import rel_close

pytestmark = pytest.mark.gpu

RTOL = 1e-12

CASES = {
    # convection: periodic sides, tanh z grid, hot bottom / cold top (probType 5), rbcType 2
    "rbc_ppn_tanh": dict(ind=(4, 4, 5), L=(2.0, 2.0, 1.0), domain="PPN", mesh="UUD", scalar=True, rbc_type=2,
                         Ra=1e5, Pr=1.0, dt=2e-3, prob_type=5, Re=1.0),
    # closed box, stable stratification (probType 6), rbcType 1
    "sst_nnn": dict(ind=(4, 4, 4), L=(1.0, 1.0, 1.0), domain="NNN", mesh="UUU", scalar=True, rbc_type=1,
                    Ra=1e3, Pr=0.7, dt=5e-4, prob_type=6, Re=1.0),
}


def _pair(case):
    from saras_amd import ts
    kw = dict(CASES[case])
    o = po.OracleTS(po.ts_params(**kw))
    g = ts.Hydro(ts.ts_desc(**kw))
    return o, g


def _cmp(g, o, wg, wo, rtol=RTOL, scale=None, exact=False, what=""):
    a, b = g.download(wg), o.array(wo)
    assert a.shape == b.shape, what
    if exact:
        assert np.array_equal(a, b), (what, float(np.max(np.abs(a - b))))
        return
    ok, err, s = rel_close(a, b, rtol, scale=scale)
    assert ok, (what, err, s)


def _random_state(o, g, seed):
    from saras_amd import ts
    rng = np.random.default_rng(seed)
    for c in range(3):
        a = o.array(c)
        a[...] = 0.1 * rng.standard_normal(a.shape)
        g.upload(c, a)
    for wo, wg in ((po.PRES, ts.P), (po.TEMP, ts.T)):
        a = o.array(wo)
        a[...] = rng.standard_normal(a.shape)
        g.upload(wg, a)


@pytest.mark.parametrize("case", list(CASES))
def test_rbc_initial_state_and_walls(case):
    from saras_amd import ts
    o, g = _pair(case)
    o.init_rbc(seed=5, amp=0.05)
    g.initRBC(seed=5, amp=0.05)
    _cmp(g, o, ts.T, po.TEMP, exact=True, what="T")
    for c in range(3):
        _cmp(g, o, c, c, exact=True, what=f"V{c}")
    _cmp(g, o, ts.P, po.PRES, exact=True, what="P")
    a = o.array(po.TEMP)
    a[...] = np.random.default_rng(2).standard_normal(a.shape)
    g.upload(ts.T, a)
    o.impose_tbc()
    g.imposeTBCs()
    _cmp(g, o, ts.T, po.TEMP, exact=True, what="T walls")
    g.close()
    o.close()


@pytest.mark.parametrize("case", list(CASES))
def test_buoyant_rhs_and_temperature_rhs(case):
    from saras_amd import ts
    o, g = _pair(case)
    _random_state(o, g, 3)
    o.build_rhs()
    g.buildRHS()
    for c, (wg, wo) in enumerate(((ts.HX, po.HX), (ts.HY, po.HY), (ts.HZ, po.HZ))):
        _cmp(g, o, wg, wo, what=f"H{c}")
    o.build_trhs()
    g.buildTRHS()
    _cmp(g, o, ts.TRHS, po.TRHS, what="tmpRHS")
    g.close()
    o.close()


@pytest.mark.parametrize("case", list(CASES))
def test_temperature_jacobi_solve(case):
    from saras_amd import ts
    o, g = _pair(case)
    o.init_rbc(seed=7, amp=0.05)
    g.initRBC(seed=7, amp=0.05)
    o.build_trhs()
    g.buildTRHS()
    it_o, it_g = o.solve_t(), g.solveT()
    assert it_o == it_g, (it_o, it_g)
    _cmp(g, o, ts.T, po.TEMP, what="T")
    g.close()
    o.close()


@pytest.mark.parametrize("case", list(CASES))
def test_scalar_time_steps(case):
    from saras_amd import ts
    o, g = _pair(case)
    o.init_rbc(seed=11, amp=0.05)
    g.initRBC(seed=11, amp=0.05)
    for step in range(4):
        o.time_advance()
        st = g.timeAdvance()
        assert [st.jacobi_iters[c] for c in range(4)] == [o.jacobi_iters(c) for c in range(4)], step
        assert st.mg_cycles == o.mg_cycles(), step
        # the velocity is the small difference of buoyancy and pressure gradient at early times:
        # its errors are measured against the size of those terms (dt Fb max|T|) or of V itself
        fb = {1: CASES[case]["Ra"] * CASES[case]["Pr"], 2: 1.0}[CASES[case]["rbc_type"]]
        vs = max(max(float(np.max(np.abs(o.array(c)))) for c in range(3)),
                 CASES[case]["dt"] * fb * float(np.max(np.abs(o.array(po.TEMP)))))
        for c in range(3):
            _cmp(g, o, c, c, scale=vs, what=f"step {step} V{c}")
        _cmp(g, o, ts.T, po.TEMP, what=f"step {step} T")
        for wg, wo in ((ts.P, po.PRES), (ts.PP, po.PCORR)):
            a, b = g.download(wg), o.array(wo)
            a, b = a - a[1:-1, 1:-1, 1:-1].mean(), b - b[1:-1, 1:-1, 1:-1].mean()
            ok, err, s = rel_close(a, b, RTOL, scale=float(np.max(np.abs(o.array(wo)))))
            assert ok, (f"step {step} pressure {wg}", err, s)
        # the energy is quadratic in V: with |dV| <= RTOL vs and |V| <= vs its error is below RTOL vs^2
        # (at the first steps E is far below vs^2, the cancellation above)
        e_o, e_g = o.energy(), g.energy()
        assert abs(e_o - e_g) <= RTOL * max(abs(e_o), vs * vs), (step, e_o, e_g, vs)
        n_o, n_g = o.nusselt(), g.nusselt()
        assert abs(n_o - n_g) <= RTOL * abs(n_o), (step, n_o, n_g)
    g.close()
    o.close()

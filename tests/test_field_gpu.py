"""Parity of the field-operator entry points (include/saras_field.h) on caller device arrays with
the oracle's restatement of lib/field, the boundary classes and the Jacobi solves.

Bar: stencil operators within 1e-12 relative (max-norm over the padded array, on the scale of the
expected array); boundary fills, which only copy, negate or set, bit for bit; Jacobi solves with
the same sweep count and the field within 1e-12.  Inputs are seeded random fields, so every
entry of every stencil takes part."""
import numpy as np
import pytest

from oracle import pyoracle as po
from helpers import rel_close

pytestmark = pytest.mark.gpu

RTOL = 1e-12

GRIDS = {
    "ppp_uniform": dict(ind=(5, 4, 4), L=(2 * np.pi,) * 3, domain="PPP", mesh="UUU"),
    "ppn_tanh": dict(ind=(4, 5, 4), L=(1.0, 1.0, 1.0), domain="PPN", mesh="UUD", beta=(1.0, 1.0, 1.2)),
    "nnn_stretched": dict(ind=(4, 3, 4), L=(1.0, 2.0, 1.5), domain="NNN", mesh="DDD", beta=(1.1, 0.7, 1.3)),
    "npp_stretched": dict(ind=(3, 4, 5), L=(1.0, 2.0, 1.0), domain="NPP", mesh="DDU", beta=(1.1, 0.7, 1.0)),
}


def _pair(grid, i_scheme=1, **kw):
    from saras_amd import field
    g = GRIDS[grid]
    o = po.OracleTS(po.ts_params(**g, **kw))
    fg = field.FieldGrid(field.fgrid_desc(ind=g["ind"], L=g["L"], mesh=g["mesh"], beta=g.get("beta", (1.0,) * 3),
                                          domain=g["domain"], i_scheme=i_scheme))
    return o, fg


def _close(a, b, what, rtol=RTOL):
    assert a.shape == b.shape, (what, a.shape, b.shape)
    ok, err, s = rel_close(a, b, rtol)
    assert ok, (what, err, s)


@pytest.mark.parametrize("grid", list(GRIDS))
def test_derivatives_every_location(grid):
    o, fg = _pair(grid)
    rng = np.random.default_rng(1)
    for loc in range(4):
        F = rng.standard_normal(o.loc_shape(loc))
        dF = fg.array(loc, F)
        out = fg.array(loc, rng.standard_normal(F.shape))   # stale contents must not survive
        for order in (1, 2):
            for d in range(3):
                (fg.calcDerivative1 if order == 1 else fg.calcDerivative2)(loc, d, dF, out)
                fg.sync()
                _close(out.download(), o.fd_deriv(loc, order, d, F), f"loc {loc} order {order} dim {d}")
    fg.close()
    o.close()


@pytest.mark.parametrize("grid", list(GRIDS))
def test_compute_diff_vector_and_scalar(grid):
    o, fg = _pair(grid)
    rng = np.random.default_rng(2)
    V = [rng.standard_normal(o.loc_shape(c + 1)) for c in range(3)]
    H = [rng.standard_normal(o.loc_shape(c + 1)) for c in range(3)]
    dV = [fg.array(c + 1, V[c]) for c in range(3)]
    dH = [fg.array(c + 1, H[c]) for c in range(3)]
    fg.vfComputeDiff(dV, dH)
    F = rng.standard_normal(o.loc_shape(0))
    HF = rng.standard_normal(F.shape)
    dF, dHF = fg.array(0, F), fg.array(0, HF)
    fg.sfComputeDiff(dF, dHF)
    fg.sync()
    for c in range(3):
        exp = H[c].copy()
        o.fd_diff(c + 1, V[c], exp)
        _close(dH[c].download(), exp, f"vfield H{c}")
    exp = HF.copy()
    o.fd_diff(0, F, exp)
    _close(dHF.download(), exp, "sfield H")
    fg.close()
    o.close()


@pytest.mark.parametrize("grid", list(GRIDS))
def test_compute_nlin_vector_and_scalar(grid):
    """vfield::computeNLin with the field advecting itself (the reference's call) and with a
    different advecting field; sfield::computeNLin of a node field."""
    o, fg = _pair(grid)
    rng = np.random.default_rng(3)
    U = [rng.standard_normal(o.loc_shape(c + 1)) for c in range(3)]
    V = [rng.standard_normal(o.loc_shape(c + 1)) for c in range(3)]
    dU = [fg.array(c + 1, U[c]) for c in range(3)]
    dV = [fg.array(c + 1, V[c]) for c in range(3)]
    for adv, dadv, what in ((U, dU, "self"), (V, dV, "other")):
        H = [rng.standard_normal(o.loc_shape(c + 1)) for c in range(3)]
        dH = [fg.array(c + 1, H[c]) for c in range(3)]
        fg.vfComputeNLin(dU, dadv, dH)
        fg.sync()
        for c in range(3):
            exp = H[c].copy()
            o.fd_nlin(c + 1, U[c], adv, exp)
            _close(dH[c].download(), exp, f"vfield nlin {what} H{c}")
    F = rng.standard_normal(o.loc_shape(0))
    HF = rng.standard_normal(F.shape)
    dF, dHF = fg.array(0, F), fg.array(0, HF)
    fg.sfComputeNLin(dV, dF, dHF)
    fg.sync()
    exp = HF.copy()
    o.fd_nlin(0, F, V, exp)
    _close(dHF.download(), exp, "sfield nlin")
    fg.close()
    o.close()


@pytest.mark.parametrize("grid", list(GRIDS))
def test_divergence_and_gradient(grid):
    o, fg = _pair(grid)
    rng = np.random.default_rng(4)
    V = [rng.standard_normal(o.loc_shape(c + 1)) for c in range(3)]
    dV = [fg.array(c + 1, V[c]) for c in range(3)]
    ddiv = fg.array(0, rng.standard_normal(o.loc_shape(0)))
    fg.divergence(dV, ddiv)
    F = rng.standard_normal(o.loc_shape(0))
    G0 = [rng.standard_normal(o.loc_shape(c + 1)) for c in range(3)]
    dG = [fg.array(c + 1, G0[c]) for c in range(3)]
    fg.gradient(fg.array(0, F), dG)
    fg.sync()
    _close(ddiv.download(), o.fd_divergence(V), "divergence")
    exp = [g.copy() for g in G0]
    o.fd_gradient(F, exp)
    for c in range(3):
        _close(dG[c].download(), exp[c], f"gradient {c}")
    fg.close()
    o.close()


def _kinds(o, rng):
    per = (o.params.xPer, o.params.yPer, o.params.zPer)
    # Dirichlet, Neumann or left alone on walls of non-periodic directions; periodic or left alone
    # (syncData only) on the others
    kinds = [int(rng.integers(1, 4)) if not per[w // 2] else int(rng.choice([0, 3])) for w in range(6)]
    values = list(rng.standard_normal(6))
    return [int(k) for k in kinds], values


@pytest.mark.parametrize("grid", list(GRIDS))
def test_bc_fill_bitwise(grid):
    """syncData plus mixes of Dirichlet, Neumann and untouched walls on every location, bit for bit."""
    from saras_amd import field
    o, fg = _pair(grid)
    rng = np.random.default_rng(5)
    for trial in range(4):
        for loc in range(4):
            kinds, values = _kinds(o, rng)
            F = rng.standard_normal(o.loc_shape(loc))
            dF = fg.array(loc, F)
            fg.imposeBC(loc, field.bc(kinds, values), dF)
            fg.sync()
            exp = F.copy()
            o.fd_bc(loc, kinds, values, exp)
            got = dF.download()
            assert np.array_equal(got, exp), (trial, loc, kinds, np.max(np.abs(got - exp)))
    fg.close()
    o.close()


def test_bc_fill_rejects_periodic_mismatch():
    from saras_amd import field, _lib
    o, fg = _pair("ppn_tanh")
    dF = fg.array(0)
    with pytest.raises(_lib.SarasError):
        fg.imposeBC(0, field.bc([1, 1, 0, 0, 1, 1]), dF)      # x is periodic on this grid
    with pytest.raises(_lib.SarasError):
        fg.imposeBC(0, field.bc([0, 0, 0, 0, 0, 0]), dF)      # z is not
    fg.close()
    o.close()


def _vbc(o, c):
    from saras_amd import field
    per = (o.params.xPer, o.params.yPer, o.params.zPer)
    kinds = [0 if per[w // 2] else 1 for w in range(6)]
    values = [1.0 if (o.params.probType == 1 and c == 0 and w == 5) else 0.0 for w in range(6)]
    return field.bc(kinds, values)


@pytest.mark.parametrize("grid,Re,dt", [("ppn_tanh", 50.0, 2e-3), ("nnn_stretched", 20.0, 5e-3),
                                        ("npp_stretched", 200.0, 1e-3)])
def test_jacobi_hydro_matches_solve_v(grid, Re, dt):
    """hydro_d3::solveVx/Vy/Vz: same sweeps, same field; the walls are imposeU/V/WBCs'."""
    o, fg = _pair(grid, Re=Re, dt=dt, prob_type=1 if grid == "nnn_stretched" else 2)
    rng = np.random.default_rng(6)
    for c in range(3):
        V, R = o.array(c), o.array(po.HX + c)
        V[...] = rng.standard_normal(V.shape) * 0.1
        R[...] = V + rng.standard_normal(V.shape) * 1e-3
        dV, dR, dG = fg.array(c + 1, V), fg.array(c + 1, R), fg.array(c + 1)
        n = fg.jacobi(c + 1, 0, dt, _vbc(o, c), dR, dV, dG, Re=Re, cn_tol=1e-6)
        n_o = o.solve_v(c)
        assert n == n_o, (c, n, n_o)
        _close(dV.download(), o.array(c), f"V{c} after {n} sweeps")
    fg.close()
    o.close()


def test_jacobi_scalar_forms_match_solve_v_and_solve_t():
    """scalar_d3::solveVx/Vy/Vz (coefficient nu) and solveT (kappa, collocated z metric at node
    indices) on the Rayleigh-Benard walls."""
    from saras_amd import field
    kw = dict(prob_type=5, scalar=True, rbc_type=2, Ra=1e4, Pr=1.0, dt=5e-2)
    o, fg = _pair("ppn_tanh", **kw)
    nu, kappa = np.sqrt(1.0 / 1e4), 1.0 / np.sqrt(1e4)
    rng = np.random.default_rng(7)
    for c in range(3):
        V, R = o.array(c), o.array(po.HX + c)
        V[...] = rng.standard_normal(V.shape) * 0.1
        R[...] = V + rng.standard_normal(V.shape) * 1e-2
        dV, dR, dG = fg.array(c + 1, V), fg.array(c + 1, R), fg.array(c + 1)
        n = fg.jacobi(c + 1, 1, kw["dt"], _vbc(o, c), dR, dV, dG, coef=nu, cn_tol=1e-6)
        assert n == o.solve_v(c), c
        _close(dV.download(), o.array(c), f"V{c}")
    T, R = o.array(po.TEMP), o.array(po.TRHS)
    T[...] = rng.standard_normal(T.shape) * 0.1
    R[...] = T + rng.standard_normal(T.shape) * 1e-2
    tb = field.bc([0, 0, 0, 0, 1, 1], [0.0, 0.0, 0.0, 0.0, 1.0, 0.0])
    dT, dR, dG = fg.array(0, T), fg.array(0, R), fg.array(0)
    n = fg.jacobi(0, 1, kw["dt"], tb, dR, dT, dG, coef=kappa, cn_tol=1e-6, z_colloc=True)
    assert n == o.solve_t()
    _close(dT.download(), o.array(po.TEMP), "T")
    fg.close()
    o.close()


def test_full_scheme_second_derivative():
    """iScheme other than 1 keeps the whole xi^2 part (derivative.cc:190-193): on a uniform grid
    the full second difference is twice the Crank-Nicolson one."""
    o, fg1 = _pair("ppp_uniform")
    _, fg0 = _pair("ppp_uniform", i_scheme=0)
    rng = np.random.default_rng(8)
    F = rng.standard_normal(o.loc_shape(0))
    dF = fg1.array(0, F)
    a, b = fg1.array(0), fg0.array(0)
    for d in range(3):
        fg1.calcDerivative2(0, d, dF, a)
        fg0.calcDerivative2(0, d, dF, b)
        fg1.sync()
        fg0.sync()
        x, y = a.download(), b.download()
        assert np.allclose(y, 2 * x, rtol=1e-13, atol=1e-12 * np.max(np.abs(y))), d
    fg1.close()
    fg0.close()
    o.close()

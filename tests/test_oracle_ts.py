"""Pins of the time-step oracle (oracle/field_oracle.c) against the reference's own unit tests
and closed forms.  CPU only.

  differ_test.cc:50-94     first derivative of the Taylor-Green field, mean |error| < 4 pi^2 h^2
  field_test.cc:55-125     divergence of the Taylor-Green velocity, max < max h
  nlin_test.cc:55-138      non-linear term of the Taylor-Green velocity, mean |error| < 2 pi max h
                           (vfield::computeNLin subtracts u.grad u from H, vfield.cc:172, so the
                           computed term is the negative of that test's analytic field)
  hydro_d3.cc:536-679      testPeriodic: after the velocity BCs every pad holds the Taylor-Green
                           value at the extrapolated coordinate, tol 1e-10
  tseries.cc:158-169       Taylor-Green energy: closed form of the discrete sum
plus exactness on quadratics, the wall conditions of hydro.cc:287-461 and the viscous decay rate."""
import numpy as np
import pytest

from oracle import pyoracle as po

TWO_PI = 2 * np.pi


def tgv_fields(o, L):
    """Taylor-Green velocity evaluated at every padded index of each component."""
    out = []
    for c in range(3):
        sh = o.shape(c)
        x = o.metric(0, c != 0, 0)[: sh[0]]
        y = o.metric(1, c != 1, 0)[: sh[1]]
        z = o.metric(2, c != 2, 0)[: sh[2]]
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        if c == 0:
            v = np.sin(TWO_PI * X / L[0]) * np.cos(TWO_PI * Y / L[1]) * np.cos(TWO_PI * Z / L[2])
        elif c == 1:
            v = -np.cos(TWO_PI * X / L[0]) * np.sin(TWO_PI * Y / L[1]) * np.cos(TWO_PI * Z / L[2])
        else:
            v = np.zeros(sh)
        out.append(v)
    return out


def core(a):
    return a[1:-1, 1:-1, 1:-1]


@pytest.fixture(scope="module")
def tgv33():
    p = po.ts_params(ind=(5, 5, 5), L=(1.0, 1.0, 1.0))
    o = po.OracleTS(p)
    o.init_tgv()
    yield o
    o.close()


def test_tgv_initial_condition_matches_closed_form(tgv33):
    V = tgv_fields(tgv33, (1.0, 1.0, 1.0))
    for c in range(3):
        assert np.allclose(tgv33.array(c), V[c], atol=1e-14)
    assert np.all(tgv33.array(po.PRES) == 1.0)


def test_first_derivative_tgv(tgv33):
    """differ_test: dF/dx of F = sin cos cos on the node grid, mean |err| over the core."""
    h = 1.0 / 32
    # derivatives of a P-shaped (node) field: load F into P
    P = tgv33.array(po.PRES)
    sh = P.shape
    x = tgv33.metric(0, True, 0)[: sh[0]]
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    saved = P.copy()
    P[...] = np.sin(TWO_PI * X) * np.cos(TWO_PI * Y) * np.cos(TWO_PI * Z)
    exact = TWO_PI * np.cos(TWO_PI * X) * np.cos(TWO_PI * Y) * np.cos(TWO_PI * Z)
    d = tgv33.deriv(po.PRES, 1, 0)
    P[...] = saved
    err = np.mean(np.abs(core(d) - core(exact)))
    assert err < 4 * np.pi ** 2 * h ** 2
    # a wrong metric or a one-sided stencil fails by far more
    assert err < 0.3 * 4 * np.pi ** 2 * h ** 2


def test_divergence_tgv(tgv33):
    """field_test: the discrete divergence of Taylor-Green vanishes (here to rounding)."""
    tgv33.divergence()
    div = tgv33.array(po.DIV)
    assert np.max(core(div)) < 1.0 / 32
    assert np.max(np.abs(core(div))) < 1e-12


def test_nonlinear_term_tgv(tgv33):
    """nlin_test: H_x = -pi sin(4 pi x) cos^2(2 pi z) on the x faces, mean |err| < 2 pi h."""
    for c in range(3):
        tgv33.array(6 + c)[...] = 0.0
    tgv33.compute_nlin()
    Hx = tgv33.array(po.HX)
    sh = Hx.shape
    x = tgv33.metric(0, False, 0)[: sh[0]]
    z = tgv33.metric(2, True, 0)[: sh[2]]
    X, _, Z = np.meshgrid(x, np.zeros(sh[1]), z, indexing="ij")
    exact = -np.pi * np.sin(2 * TWO_PI * X) * np.cos(TWO_PI * Z) ** 2
    err = np.mean(np.abs(core(Hx) - core(exact)))
    assert err < TWO_PI / 32
    assert err < 0.2 * TWO_PI / 32
    Hz = tgv33.array(po.HZ)
    assert np.max(np.abs(core(Hz))) < 1e-14      # w = 0 and dw = 0


def test_hydro_test_periodic_pads():
    """hydro_d3::testPeriodic: pads refilled by the BCs equal Taylor-Green at the extrapolated
    coordinates (uniform periodic grid: the padded coordinate arrays already extrapolate)."""
    p = po.ts_params(ind=(4, 4, 4), L=(TWO_PI,) * 3)
    o = po.OracleTS(p)
    o.init_tgv()
    V = tgv_fields(o, (TWO_PI,) * 3)
    for c in range(3):
        a = o.array(c)
        a[0, :, :] = 7.0
        a[-1, :, :] = 7.0
        a[:, 0, :] = 7.0
        a[:, -1, :] = 7.0
        a[:, :, 0] = 7.0
        a[:, :, -1] = 7.0
        o.impose_vbc(c)
        assert np.max(np.abs(a - V[c])) < 1e-10, c
    o.close()


def test_energy_closed_form():
    """E = sum over all core nodes (duplicated periodic nodes included) of the squared face
    averages, volume-weighted, no 1/2: for Taylor-Green on N nodes,
    E = 2 a b^2 cos^2(pi/(N-1)) / N^3, a = (N-1)/2, b = a + 1."""
    for ind in (4, 5):
        p = po.ts_params(ind=(ind,) * 3, L=(TWO_PI,) * 3)
        o = po.OracleTS(p)
        o.init_tgv()
        N = 2 ** ind + 1
        a = (N - 1) / 2
        b = a + 1
        expect = 2 * a * b * b * np.cos(np.pi / (N - 1)) ** 2 / N ** 3
        assert abs(o.energy() - expect) < 1e-11 * expect
        o.close()


@pytest.mark.parametrize("mesh", ["UUU", "UUD"])
def test_derivatives_exact_on_quadratics(mesh):
    """Central stencils are exact for quadratics in the computational coordinate:
    F = xi^2 gives d1 = 2 xi xi_x and d2 = xi_xx 2 xi + 0.5 xi_x^2 2 (the CN half on the second part)."""
    p = po.ts_params(ind=(4, 4, 4), L=(1.0, 1.0, 1.0), domain="PPN", mesh=mesh)
    o = po.OracleTS(p)
    P = o.array(po.PRES)
    n = P.shape[2] - 2
    xi = np.arange(-1, n + 1) / (n - 1)
    P[...] = (xi ** 2)[None, None, :]
    d1 = o.deriv(po.PRES, 1, 2)
    d2 = o.deriv(po.PRES, 2, 2)
    m1 = o.metric(2, True, 1)
    mxx = o.metric(2, True, 2)
    mx2 = o.metric(2, True, 3)
    for k in range(n):
        assert np.allclose(d1[:, :, k + 1], 2 * xi[k + 1] * m1[k + 1], rtol=1e-12, atol=1e-12)
        assert np.allclose(d2[:, :, k + 1], mxx[k + 1] * 2 * xi[k + 1] + 0.5 * mx2[k + 1] * 2, rtol=1e-11, atol=1e-11)
    # outside the core range along the direction the output stays zero
    assert np.all(d1[:, :, 0] == 0) and np.all(d1[:, :, -1] == 0)
    o.close()


def test_gradient_exact_on_linear_pressure():
    """sfield::gradient of P = x is the face metric times h / h = 1 on a uniform grid."""
    p = po.ts_params(ind=(4, 4, 4), L=(2.0, 2.0, 2.0))
    o = po.OracleTS(p)
    P = o.array(po.PRES)
    x = o.metric(0, True, 0)[: P.shape[0]]
    P[...] = x[:, None, None]
    o.gradient(False)
    G = o.array(9)
    assert np.allclose(core(G), 1.0, rtol=1e-13)
    o.close()


def test_wall_conditions_probtype2():
    """hydro::initVBC walls (probType 2): walls normal to a component hold 2*0 - interior (no
    penetration on the face grid), the others 0 at the wall node (no slip)."""
    p = po.ts_params(ind=(4, 4, 4), L=(1.0, 1.0, 1.0), domain="NNN")
    o = po.OracleTS(p)
    rng = np.random.default_rng(3)
    for c in range(3):
        a = o.array(c)
        a[...] = rng.standard_normal(a.shape)
        o.impose_vbc(c)
    Vx, Vy, Vz = o.array(0), o.array(1), o.array(2)
    # Vx: x faces; ghost faces antisymmetric, interior columns only (corners follow the order)
    assert np.allclose(Vx[0, 2:-2, 2:-2], -Vx[1, 2:-2, 2:-2])
    assert np.allclose(Vx[-1, 2:-2, 2:-2], -Vx[-2, 2:-2, 2:-2])
    # Vx at the y and z wall nodes (index 0 and N-1 of the core) is zero
    assert np.all(Vx[:, 1, :] == 0) and np.all(Vx[:, -2, :] == 0)
    assert np.all(Vx[:, :, 1] == 0) and np.all(Vx[:, :, -2] == 0)
    assert np.all(Vy[1, :, :] == 0) and np.all(Vy[-2, :, :] == 0)
    assert np.allclose(Vz[2:-2, 2:-2, 0], -Vz[2:-2, 2:-2, 1])
    o.close()


def test_jacobi_fixed_point():
    """After solveVx the bulk satisfies |u - (dt/2Re) lap u - rhs| < cnTolerance, with the Laplacian
    written here from the metric terms (hydro_d3.cc:396-410)."""
    p = po.ts_params(ind=(4, 4, 4), L=(TWO_PI,) * 3, Re=10.0, dt=1e-2, cn_tol=1e-10)
    o = po.OracleTS(p)
    o.init_tgv()
    o.build_rhs()
    it = o.solve_v(0)
    assert it > 1
    V = o.array(0)
    R = o.array(po.HX)
    h = 1.0 / 16
    mx = o.metric(0, False, 3)[1:-1]
    my = o.metric(1, True, 3)[1:-1]
    mz = o.metric(2, True, 3)[1:-1]
    c = V[1:-1, 1:-1, 1:-1]
    lap = (mx[:, None, None] * (V[2:, 1:-1, 1:-1] - 2 * c + V[:-2, 1:-1, 1:-1]) +
           my[None, :, None] * (V[1:-1, 2:, 1:-1] - 2 * c + V[1:-1, :-2, 1:-1]) +
           mz[None, None, :] * (V[1:-1, 1:-1, 2:] - 2 * c + V[1:-1, 1:-1, :-2])) / h ** 2
    res = np.abs(c - 0.5 * p.dt / p.Re * lap - core(R))
    assert res.max() < 1e-10
    o.close()


def test_tgv_viscous_decay_rate():
    """Early Taylor-Green decay: dE/dt = -6 E / Re in the continuum (|k|^2 = 3 per component with
    L = 2 pi); the discrete rate agrees within 3 % after a few steps (a dropped or doubled viscous
    term, or a wrong sign, fails)."""
    p = po.ts_params(ind=(5, 5, 5), L=(TWO_PI,) * 3, Re=100.0, dt=1e-3)
    o = po.OracleTS(p)
    o.init_tgv()
    e0 = o.energy()
    steps = 4
    for _ in range(steps):
        o.time_advance()
    e1 = o.energy()
    rate = -np.log(e1 / e0) / (steps * p.dt)
    assert abs(rate - 6.0 / p.Re) < 0.03 * 6.0 / p.Re, rate
    # the projection keeps the velocity discretely divergence-free
    assert abs(o.max_div()) < 1e-5
    o.close()

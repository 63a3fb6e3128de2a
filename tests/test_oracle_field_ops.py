"""Pins of the oracle's field operators on caller arrays (or_fd_*, oracle/field_oracle.c), which
the field-operator entry points of include/saras_field.h are checked against.  CPU only.

Each operator is pinned by what the mathematics fixes on the discrete grid, independently of the
stencil code: central differences are exact on polynomials of degree <= 2 in the computational
coordinates (xi, eta, zeta), so derivative-based operators applied to such fields reduce to
closed forms in the metric arrays (themselves pinned against the tanh map in test_oracle_mg);
interpolation averages are exact on constants; the boundary routine reproduces the wall rules of
imposeU/V/WBCs and imposeTBCs, which test_oracle_ts / test_oracle_scalar pin."""
import numpy as np
import pytest

from oracle import pyoracle as po

GRIDS = {
    "ppp_uniform": dict(ind=(4, 4, 4), L=(2 * np.pi,) * 3, domain="PPP", mesh="UUU"),
    "ppn_tanh": dict(ind=(4, 3, 4), L=(1.0, 1.0, 1.0), domain="PPN", mesh="UUD", beta=(1.0, 1.0, 1.2)),
    "nnn_stretched": dict(ind=(3, 4, 3), L=(1.0, 2.0, 1.5), domain="NNN", mesh="DDD", beta=(1.1, 0.7, 1.3)),
}


def _oracle(grid, **kw):
    return po.OracleTS(po.ts_params(**GRIDS[grid], **kw))


def _coords(o, loc):
    """Computational coordinates of every padded entry of a location: xi = h * index, index
    -1..n along each direction (nodes at i h, faces at (i + 1/2) h; only differences of xi enter
    the stencils, so the face offset is immaterial for the exactness argument)."""
    sh = o.loc_shape(loc)
    h = [1.0 / ((1 << n)) for n in (o.params.xInd, o.params.yInd, o.params.zInd)]
    ax = [h[d] * (np.arange(sh[d]) - 1) for d in range(3)]
    return np.meshgrid(*ax, indexing="ij"), h


def _metric(o, loc, d, q):
    """metric quantity q along d for location loc, indexed by padded position (length n + 2)"""
    node = loc == 0 or loc - 1 != d
    m = o.metric(d, node, q)
    n = o.loc_shape(loc)[d]
    return m[:n]


def _bcast(v, d):
    s = [1, 1, 1]
    s[d] = -1
    return v.reshape(s)


@pytest.mark.parametrize("grid", list(GRIDS))
@pytest.mark.parametrize("loc", [0, 1, 2, 3])
def test_diff_exact_on_quadratics(grid, loc):
    """computeDiff of F = sum_d c_d xi_d^2 adds sum_d (xi_dd 2 c_d xi_d + 0.5 xi_d^2 2 c_d) on the core
    (iScheme 1) and leaves the pads of H alone."""
    o = _oracle(grid)
    X, _ = _coords(o, loc)
    c = (1.0, -2.0, 0.5)
    F = sum(c[d] * X[d] ** 2 for d in range(3))
    H0 = np.random.default_rng(1).standard_normal(F.shape)
    H = H0.copy()
    o.fd_diff(loc, F, H)
    exp = H0.copy()
    core = (slice(1, -1),) * 3
    add = sum(_bcast(_metric(o, loc, d, 2), d) * 2 * c[d] * X[d] + 0.5 * _bcast(_metric(o, loc, d, 3), d) * 2 * c[d]
              for d in range(3))
    exp[core] += add[core]
    assert np.allclose(H, exp, rtol=1e-11, atol=1e-9 * np.max(np.abs(add[core])))
    o.close()


@pytest.mark.parametrize("grid", list(GRIDS))
@pytest.mark.parametrize("loc", [0, 1, 2, 3])
def test_nlin_exact_on_linear_field_constant_velocity(grid, loc):
    """computeNLin of F = sum_d a_d xi_d advected by constant V_d = U_d subtracts
    sum_d U_d xi_x,d a_d on the core (every interpolation slice average of a constant is that
    constant, the central difference of a linear function is its slope)."""
    o = _oracle(grid)
    X, _ = _coords(o, loc)
    a = (0.7, -1.3, 2.1)
    U = (1.5, -0.25, 0.8)
    F = sum(a[d] * X[d] for d in range(3))
    V = [np.full(o.loc_shape(c + 1), U[c]) for c in range(3)]
    H = np.zeros(F.shape)
    o.fd_nlin(loc, F, V, H)
    exp = np.zeros(F.shape)
    core = (slice(1, -1),) * 3
    sub = sum(U[d] * _bcast(_metric(o, loc, d, 1), d) * a[d] for d in range(3)) * np.ones(F.shape)
    exp[core] -= sub[core]
    assert np.allclose(H, exp, rtol=1e-11, atol=1e-12 * np.max(np.abs(sub)))
    o.close()


@pytest.mark.parametrize("grid", list(GRIDS))
def test_gradient_and_divergence_exact_on_linear_fields(grid):
    """gradient of F = sum_d a_d xi_d is xi_c(face) a_c on each face core; divergence of V_c = b_c
    xi_c is sum_c xi_c(node) b_c on the node core and 0 on its pads."""
    o = _oracle(grid)
    X, _ = _coords(o, 0)
    a = (0.3, -0.9, 1.7)
    F = sum(a[d] * X[d] for d in range(3))
    G = [np.full(o.loc_shape(c + 1), 7.0) for c in range(3)]
    o.fd_gradient(F, G)
    for c in range(3):
        exp = np.full(G[c].shape, 7.0)
        core = (slice(1, -1),) * 3
        val = _bcast(_metric(o, c + 1, c, 1), c) * a[c] * np.ones(G[c].shape)
        exp[core] = val[core]
        assert np.allclose(G[c], exp, rtol=1e-12, atol=1e-13), c
    b = (1.1, 0.4, -2.2)
    V = []
    for c in range(3):
        Xc, _ = _coords(o, c + 1)
        V.append(b[c] * Xc[c])
    div = o.fd_divergence(V)
    exp = np.zeros(div.shape)
    core = (slice(1, -1),) * 3
    val = sum(_bcast(_metric(o, 0, c, 1), c) * b[c] for c in range(3)) * np.ones(div.shape)
    exp[core] = val[core]
    assert np.allclose(div, exp, rtol=1e-12, atol=1e-13)
    o.close()


@pytest.mark.parametrize("grid", list(GRIDS))
@pytest.mark.parametrize("loc", [0, 1, 2, 3])
@pytest.mark.parametrize("order", [1, 2])
def test_deriv_on_caller_array_equals_field_derivative(grid, loc, order):
    """or_fd_deriv on a caller copy of a field equals or_ts_deriv on the field itself (whose
    Taylor-Green closed forms and quadratic exactness test_oracle_ts pins)."""
    o = _oracle(grid)
    rng = np.random.default_rng(loc + 10 * order)
    which = po.PRES if loc == 0 else loc - 1
    A = o.array(which)
    A[...] = rng.standard_normal(A.shape)
    for d in range(3):
        assert np.array_equal(o.fd_deriv(loc, order, d, A.copy()), o.deriv(which, order, d))
    o.close()


def _vbc_kinds(o, c):
    per = (o.params.xPer, o.params.yPer, o.params.zPer)
    kinds, vals = [], []
    for w in range(6):
        d = w // 2
        kinds.append(0 if per[d] else 1)
        vals.append(1.0 if (o.params.probType == 1 and c == 0 and w == 5) else 0.0)
    return kinds, vals


@pytest.mark.parametrize("grid,prob", [("ppp_uniform", 2), ("ppn_tanh", 2), ("nnn_stretched", 1)])
def test_bc_fill_reproduces_velocity_and_temperature_walls(grid, prob):
    """The generic boundary routine with no-slip / no-penetration Dirichlet walls (the lid of
    probType 1 at 1) equals imposeU/V/WBCs; with adiabatic side walls, hot bottom and cold top it
    equals imposeTBCs (probType 5), bit for bit."""
    o = _oracle(grid, prob_type=prob)
    rng = np.random.default_rng(3)
    for c in range(3):
        A = o.array(c)
        A[...] = rng.standard_normal(A.shape)
        F = A.copy()
        kinds, vals = _vbc_kinds(o, c)
        o.fd_bc(c + 1, kinds, vals, F)
        o.impose_vbc(c)
        assert np.array_equal(F, o.array(c)), c
    o.close()
    o = _oracle(grid, prob_type=5, scalar=True)
    per = (o.params.xPer, o.params.yPer, o.params.zPer)
    T = o.array(po.TEMP)
    T[...] = rng.standard_normal(T.shape)
    F = T.copy()
    kinds = [0 if per[w // 2] else (2 if w < 4 else 1) for w in range(6)]
    vals = [0.0, 0.0, 0.0, 0.0, 1.0, 0.0]
    o.fd_bc(0, kinds, vals, F)
    o.impose_tbc()
    assert np.array_equal(F, o.array(po.TEMP))
    o.close()


def test_bc_fill_wall_rules():
    """Neumann walls copy the first interior layer beyond the wall slice; Dirichlet on a face
    direction makes the mean of the image and the first face equal the value; Dirichlet on a node
    direction sets the wall slice (the boundary node just outside the bulk)."""
    o = _oracle("nnn_stretched")
    rng = np.random.default_rng(5)
    # Vx: faces in x (FC walls), nodes in y and z (CC walls)
    F = rng.standard_normal(o.loc_shape(1))
    kinds = [1, 2, 2, 1, 1, 2]
    vals = [0.5, 0.0, 0.0, -0.25, 1.5, 0.0]
    o.fd_bc(1, kinds, vals, F)
    n = o.loc_shape(1)
    # x- (FC Dirichlet): (F[-1] + F[0]) / 2 = 0.5 on the plane of the wall
    assert np.allclose((F[0, 2:-2, 2:-2] + F[1, 2:-2, 2:-2]) / 2, 0.5, atol=1e-14)
    # x+ (FC Neumann): F[n] = F[n - 1]
    assert np.array_equal(F[-1, 2:-2, 2:-2], F[-2, 2:-2, 2:-2])
    # y- (CC Neumann): the bulk drops the end node j = 0, whose value becomes that of j = 1
    assert np.array_equal(F[2:-2, 1, 2:-2], F[2:-2, 2, 2:-2])
    # y+ (CC Dirichlet): the end node j = n - 1 holds the value
    assert np.all(F[2:-2, -2, 2:-2] == -0.25)
    # z- (CC Dirichlet) and z+ (CC Neumann)
    assert np.all(F[:, :, 1] == 1.5)
    assert np.array_equal(F[:, :, -2], F[:, :, -3])
    assert n == F.shape
    o.close()

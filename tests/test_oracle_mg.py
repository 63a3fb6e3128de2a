"""Pins of the CPU oracle's multigrid (oracle/mg_oracle.c) against the reference's known-answer
tests, closed forms and brute force.  CPU only.

Reference KATs restated here (lib/poisson/poisson_d3.cc):
  testProlong 689-719, testTransfer 722-777, testPeriodic 780-857.
As committed, those three functions index a level outside the hierarchy after their loops
(vLevel-1 = -1 in testProlong, vcDepth+1 in the other two) and testProlong fills the coarse level
with the fine-index formula; the tests below keep their data generators and tolerance (1e-10,
src/tests/units/poisson_test.cc:73) and check what they are written to check, level by level.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from helpers import Problem, r2_exact, r2_walls, sin_problem


def lin(n, shift=1000.0, scale=1):
    i, j, k = (np.arange(m) * scale for m in n)
    return shift + 100.0 * i[:, None, None] + 10.0 * j[None, :, None] + k[None, None, :]


# ------------------------------------------------------------------------------------------------
# grid metrics (lib/grid/grid.cc)
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("beta,L", [(1.0, 1.0), (2.0, 1.0), (1.3, 2.5)])
def test_tanh_metrics_are_derivatives_of_the_map(beta, L):
    n = 65
    ax = po.grid_axis(n, L, 2, beta)
    xi = np.arange(n) / (n - 1)
    x = L * (1.0 - np.tanh(beta * (1.0 - 2.0 * xi)) / np.tanh(beta)) / 2.0
    assert np.allclose(ax["x"][1:-1], x, rtol=0, atol=1e-14 * L)
    # xi_x = 1 / (dx/dxi): dx/dxi = L beta sech^2(beta(1-2xi)) / tanh(beta)
    dxdxi = L * beta / np.cosh(beta * (1.0 - 2.0 * xi)) ** 2 / np.tanh(beta)
    assert np.allclose(ax["xi_x"][1:-1], 1.0 / dxdxi, rtol=1e-12, atol=0)
    assert np.allclose(ax["xix2"][1:-1], ax["xi_x"][1:-1] ** 2, rtol=1e-15, atol=0)
    # xi_xx = d(xi_x)/dx, checked with a fine central difference of the closed-form xi_x(x)
    eps = 1e-6
    def xi_x_of(xv):
        s = 1.0 - 2.0 * xv / L
        return np.tanh(beta) / (beta * L * (1.0 - (s * np.tanh(beta)) ** 2))
    fd = (xi_x_of(x + eps) - xi_x_of(x - eps)) / (2 * eps)
    assert np.allclose(ax["xixx"][1:-1], fd, rtol=1e-6, atol=1e-6 * np.max(np.abs(fd)))
    # pads keep the uniform values (grid.cc:373-420 before createTanHypGrid)
    assert ax["xi_x"][0] == 1.0 / L and ax["xixx"][0] == 0.0 and ax["xix2"][-1] == (1.0 / L) ** 2


def test_uniform_and_face_metrics():
    n, L = 17, 2.0 * np.pi
    ax = po.grid_axis(n, L, 1, 1.0)
    assert np.allclose(ax["x"], L * np.arange(-1, n + 1) / (n - 1))
    assert np.all(ax["xixx"] == 0.0) and np.allclose(ax["xix2"], 1.0 / L ** 2)
    fc = po.grid_axis(n, L, 1, 1.0, kind="colloc")
    assert np.allclose(fc["x"], L * (np.arange(-1, n) + 0.5) / (n - 1))
    ft = po.grid_axis(n, 1.0, 2, 1.0, kind="colloc")
    nt = po.grid_axis(n, 1.0, 2, 1.0)
    # faces lie between their two nodes
    assert np.all(ft["x"][1:-1] > nt["x"][1:-2]) and np.all(ft["x"][1:-1] < nt["x"][2:-1])


# ------------------------------------------------------------------------------------------------
# reference KATs
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("ind", [(3, 3, 3), (4, 3, 5), (5, 4, 3)])
def test_prolong_reproduces_trilinear_data(ind):
    """testProlong: trilinear prolongation reproduces 1000 + 100 i + 10 j + k exactly."""
    pb = Problem(ind=ind, domain="NNN", mesh="UUU")
    o = pb.oracle()
    for lev in range(1, pb.depth + 1):
        nc = o.dims(lev)
        nf = o.dims(lev - 1)
        o.level = lev
        P = o.array(po.P, lev)
        P[...] = 0.0
        P[1:-1, 1:-1, 1:-1] = lin(nc, scale=2)   # coarse node I sits on fine node 2I
        o.prolong()
        assert o.level == lev - 1
        got = o.array(po.P, lev - 1)
        assert np.max(np.abs(got[1:-1, 1:-1, 1:-1] - lin(nf))) < 1e-10
        # prolong clears the fine level's pads (poisson_d3.cc:307-309)
        assert np.all(got[0] == 0) and np.all(got[:, :, -1] == 0)


@pytest.mark.parametrize("domain", ["PPN", "PNN", "NPP"])
def test_transfer_pads(domain):
    """testTransfer: after updatePads the x/y pads hold 1000 + 100 (M-1) + 10 j + k etc."""
    pb = Problem(ind=(4, 4, 4), domain=domain, mesh="UUU")
    o = pb.oracle()
    for lev in range(pb.depth + 1):
        nx, ny, nz = o.dims(lev)
        o.level = lev
        P = o.array(po.P, lev)
        P[...] = 0.0
        P[1:-1, 1:-1, 1:-1] = lin((nx, ny, nz))
        o.update_pads(po.P)
        Mx, My = nx - 1, ny - 1
        j, k = np.arange(ny), np.arange(nz)
        i = np.arange(nx)
        if pb.per[0]:
            exp_lo = 1000 + (Mx - 1) * 100 + 10 * j[:, None] + k[None, :]
            exp_hi = 1000 + 100 + 10 * j[:, None] + k[None, :]
            assert np.max(np.abs(P[0, 1:-1, 1:-1] - exp_lo)) < 1e-10
            assert np.max(np.abs(P[-1, 1:-1, 1:-1] - exp_hi)) < 1e-10
        else:
            assert np.all(P[0] == 0) and np.all(P[-1] == 0)
        if pb.per[1]:
            exp_lo = 1000 + 100 * i[:, None] + (My - 1) * 10 + k[None, :]
            exp_hi = 1000 + 100 * i[:, None] + 10 + k[None, :]
            assert np.max(np.abs(P[1:-1, 0, 1:-1] - exp_lo)) < 1e-10
            assert np.max(np.abs(P[1:-1, -1, 1:-1] - exp_hi)) < 1e-10
        # z pads are never exchanged
        assert np.all(P[:, :, 0] == 0) and np.all(P[:, :, -1] == 0)


@pytest.mark.parametrize("domain", ["PPP", "PPN"])
def test_periodic_ghosts_match_the_analytic_field(domain):
    """testPeriodic: imposeBC ghosts of sin(2 pi x) cos(2 pi y) cos(2 pi z) equal the field at the
    ghost coordinate x(0) - (x(stride) - x(0)) of every level (uniform grid)."""
    pb = Problem(ind=(5, 4, 5), domain=domain, mesh="UUU")
    o = pb.oracle()
    xs = [np.arange(-1, k + 1) / (k - 1) for k in pb.n]
    for lev in range(pb.depth + 1):
        s = 1 << lev
        o.level = lev
        o.set_zero_bc(1)
        cx = [x[1::s][: o.dims(lev)[d]] for d, x in enumerate(xs)]
        h = [s / (k - 1) for k in pb.n]
        g = [np.concatenate([[c[0] - h[d]], c, [c[-1] + h[d]]]) for d, c in enumerate(cx)]
        X, Y, Z = np.meshgrid(*g, indexing="ij")
        F = np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) * np.cos(2 * np.pi * Z)
        P = o.array(po.P, lev)
        P[...] = 0.0
        P[1:-1, 1:-1, 1:-1] = F[1:-1, 1:-1, 1:-1]
        o.impose_bc()
        per = [pb.per[0], pb.per[1], pb.per[2]]
        for d in range(3):
            if not per[d]:
                continue
            for side in (0, -1):
                sl = [slice(1, -1)] * 3
                sl[d] = side
                assert np.max(np.abs(P[tuple(sl)] - F[tuple(sl)])) < 1e-10


# ------------------------------------------------------------------------------------------------
# restriction / prolongation structure
# ------------------------------------------------------------------------------------------------
def _apply_columns(o, lev, which_in, op, which_out, out_lev):
    nf = o.dims(lev)
    cols = []
    for idx in np.ndindex(*nf):
        o.level = lev
        A = o.array(which_in, lev)
        A[...] = 0.0
        A[tuple(i + 1 for i in idx)] = 1.0
        op()
        cols.append(o.array(which_out, out_lev)[1:-1, 1:-1, 1:-1].ravel().copy())
    return np.array(cols).T


def test_restriction_is_scaled_transpose_of_prolongation():
    """Full weighting R = P^T / 8 on interior coarse nodes; both read from the oracle by brute force."""
    pb = Problem(ind=(3, 3, 3), domain="NNN", mesh="UUU", depth=1)
    o = pb.oracle()
    R = _apply_columns(o, 0, po.T, o.coarsen, po.R, 1)        # coarse x fine
    Pm = _apply_columns(o, 1, po.P, o.prolong, po.P, 0)       # fine x coarse
    nc = o.dims(1)
    inner = [I for I, idx in enumerate(np.ndindex(*nc)) if all(0 < q < m - 1 for q, m in zip(idx, nc))]
    assert len(inner) > 0
    assert np.max(np.abs(R[inner] - Pm.T[inner] / 8.0)) < 1e-15
    # weights sum to one (restriction of a constant is that constant in the interior)
    assert np.allclose(R[inner].sum(axis=1), 1.0)


def test_restriction_reproduces_linear_data_inside():
    pb = Problem(ind=(4, 4, 4), domain="NNN", mesh="UUU", depth=1)
    o = pb.oracle()
    o.level = 0
    T = o.array(po.T, 0)
    n = o.dims(0)
    T[...] = 0.0
    T[1:-1, 1:-1, 1:-1] = lin(n)
    o.coarsen()
    got = o.array(po.R, 1)[2:-2, 2:-2, 2:-2]
    exp = lin(o.dims(1), scale=2)[1:-1, 1:-1, 1:-1]
    assert np.max(np.abs(got - exp)) < 1e-10


# ------------------------------------------------------------------------------------------------
# the discrete operator
# ------------------------------------------------------------------------------------------------
def test_laplacian_exact_on_r2_over_6():
    """TEST_POISSON: with u = r^2/6 and Dirichlet ghosts from the same formula, f - L u = 1 - 1 = 0."""
    n = (17, 17, 17)
    pb = Problem(ind=(4, 4, 4), domain="NNN", mesh="UUU", test_poisson=True, walls=r2_walls(n))
    o = pb.oracle()
    o.level = 0
    o.set_zero_bc(0)
    P = o.array(po.P, 0)
    P[1:-1, 1:-1, 1:-1] = r2_exact(n)
    o.impose_bc()
    R = o.array(po.R, 0)
    R[1:-1, 1:-1, 1:-1] = 1.0
    o.compute_residual()
    assert np.max(np.abs(o.array(po.T, 0)[1:-1, 1:-1, 1:-1])) < 1e-10


def test_tanh_operator_exact_on_quadratics_in_xi():
    """On a stretched direction L u = xi_x^2 g''(xi) + xi_xx g'(xi) holds exactly for g quadratic."""
    pb = Problem(ind=(4, 4, 5), domain="PPN", mesh="UUD")
    o = pb.oracle()
    nz = pb.n[2]
    ax = po.grid_axis(nz, 1.0, 2, 1.0)
    xi = np.arange(-1, nz + 1) / (nz - 1)
    a, b = 0.7, -0.3
    g = a * xi * xi + b * xi
    o.level = 0
    P = o.array(po.P, 0)
    P[...] = g[None, None, :]
    R = o.array(po.R, 0)
    R[...] = 0.0
    o.compute_residual()
    exp = -(ax["xix2"][1:-1] * 2 * a + ax["xixx"][1:-1] * (2 * a * xi[1:-1] + b))
    got = o.array(po.T, 0)[1:-1, 1:-1, 1:-1]
    assert np.max(np.abs(got - exp[None, None, :])) < 1e-9 * np.max(np.abs(exp))


def _dense_operator(o, lev=0, zero_bc=1):
    """Columns of A (L u at core nodes) for homogeneous ghosts: zeroBC = 1 gives the antisymmetric
    (or mirror / periodic) ghosts, zeroBC = 0 with zero wall planes gives fixed zero ghosts."""
    n = o.dims(lev)
    cols = []
    for idx in np.ndindex(*n):
        o.level = lev
        o.set_zero_bc(zero_bc)
        P = o.array(po.P, lev)
        P[...] = 0.0
        P[tuple(i + 1 for i in idx)] = 1.0
        o.impose_bc()
        o.array(po.R, lev)[...] = 0.0
        o.compute_residual()
        cols.append(-o.array(po.T, lev)[1:-1, 1:-1, 1:-1].ravel().copy())
    return np.array(cols).T


@pytest.mark.parametrize("smoother", [po.SMOOTH_LEX, po.SMOOTH_RB, po.SMOOTH_JACOBI])
def test_one_sweep_identities(smoother):
    """Jacobi: u1 = u0 + D^-1 (f - A u0).  Red-black: the black half sweep comes last and reads only
    red neighbours, so every black node has zero residual after the sweep.  Lexicographic: the first
    node is updated from old values only.  All orders converge to the dense solution A^-1 f."""
    pb = Problem(ind=(2, 2, 3), domain="NNN", mesh="UUD", test_poisson=True, depth=1)
    o = pb.oracle(smoother=smoother)
    A = _dense_operator(o)
    n = o.dims(0)
    rng = np.random.default_rng(3)
    u0 = rng.standard_normal(n)
    f = rng.standard_normal(n)
    o.level = 0
    o.set_zero_bc(1)
    o.array(po.P, 0)[...] = 0.0
    o.array(po.P, 0)[1:-1, 1:-1, 1:-1] = u0
    o.array(po.R, 0)[1:-1, 1:-1, 1:-1] = f
    o.smooth(1)
    u1 = o.array(po.P, 0)[1:-1, 1:-1, 1:-1].ravel()
    fr, u0r = f.ravel(), u0.ravel()
    if smoother == po.SMOOTH_JACOBI:
        D = np.diag(A)
        exp = u0r + (fr - A @ u0r) / D
        assert np.max(np.abs(u1 - exp)) < 1e-12 * np.max(np.abs(exp))
    elif smoother == po.SMOOTH_RB:
        black = [q for q, idx in enumerate(np.ndindex(*n)) if sum(idx) % 2 == 1]
        r = fr - A @ u1
        assert np.max(np.abs(r[black])) < 1e-10 * np.max(np.abs(A)) * np.max(np.abs(u1))
        red = [q for q, idx in enumerate(np.ndindex(*n)) if sum(idx) % 2 == 0]
        assert np.max(np.abs(r[red])) > 1e-6
    else:
        exp0 = (fr[0] - (A[0] @ u0r - A[0, 0] * u0r[0])) / A[0, 0]
        assert abs(u1[0] - exp0) < 1e-12 * max(1.0, abs(exp0))
    # fixed point: many sweeps converge to A^-1 f
    o.smooth(400)
    u = o.array(po.P, 0)[1:-1, 1:-1, 1:-1].ravel()
    assert np.max(np.abs(u - np.linalg.solve(A, f.ravel()))) < 1e-9


def test_vcycles_converge_to_dense_solution():
    pb = Problem(ind=(3, 3, 3), domain="NNN", mesh="UUD", test_poisson=True, vc_count=30,
                 semantics=po.SEM_REFERENCE)
    o = pb.oracle()
    A = _dense_operator(o, zero_bc=0)   # level 0 of a V-cycle uses the (zero) wall planes
    f = pb.random_rhs(1)
    p, hist = o.solve(np.zeros(pb.shape), f)
    x = np.linalg.solve(A, f[1:-1, 1:-1, 1:-1].ravel())
    assert np.max(np.abs(p[1:-1, 1:-1, 1:-1].ravel() - x)) < 1e-9 * np.max(np.abs(x))
    assert len(hist) == 30          # TEST_POISSON: no early stop


def test_r2_test_poisson_case():
    """tests/testPoisson.sh restated on one rank at 33^3: RHS = 1, walls r^2/6, 25 V-cycles."""
    n = (33, 33, 33)
    pb = Problem(ind=(5, 5, 5), domain="NNN", mesh="UUU", test_poisson=True, walls=r2_walls(n), vc_count=25,
                 depth=4, semantics=po.SEM_REFERENCE)
    for sm in (po.SMOOTH_LEX, po.SMOOTH_RB):
        o = pb.oracle(smoother=sm)
        f = np.zeros(pb.shape)
        f[1:-1, 1:-1, 1:-1] = 1.0
        p, hist = o.solve(np.zeros(pb.shape), f)
        assert len(hist) == 25
        assert np.max(np.abs(p[1:-1, 1:-1, 1:-1] - r2_exact(n))) < 1e-8


@pytest.mark.parametrize("ind", [(4, 4, 4), (5, 5, 5)])
def test_c1_sin_manufactured_error_is_second_order(ind):
    """Config C1 (33^3, V(2,2), relative residual 1e-6, reference lexicographic GS)."""
    n = [(1 << i) + 1 for i in ind]
    f, u, walls = sin_problem(n)
    pb = Problem(ind=ind, domain="NNN", mesh="UUU", test_poisson=True, walls=walls, semantics=po.SEM_PROJECTED,
                 rel_tol=1e-6)
    o = pb.oracle(smoother=po.SMOOTH_LEX)
    p, hist = o.solve(np.zeros(pb.shape), f)
    assert hist[-1] < 1e-6
    err = np.max(np.abs(p[1:-1, 1:-1, 1:-1] - u[1:-1, 1:-1, 1:-1]))
    h = 1.0 / (n[0] - 1)
    # truncation error of the 7-point scheme: |u - u_h| <= (pi^2 h^2 / 4) max|u| for this mode
    assert err < (np.pi ** 2) * h * h / 4.0
    assert err > (np.pi ** 2) * h * h / 64.0


def test_compute_error_norms():
    pb = Problem(ind=(3, 3, 3), domain="PPN", mesh="UUD")
    o = pb.oracle()
    rng = np.random.default_rng(5)
    o.level = 0
    P, R = o.array(po.P, 0), o.array(po.R, 0)
    P[...] = rng.standard_normal(P.shape)
    R[...] = rng.standard_normal(R.shape)
    o.array(po.T, 0)[...] = 0.0
    o.compute_residual()
    r = o.array(po.T, 0)[1:-1, 1:-1, 1:-1]
    fc = R[1:-1, 1:-1, 1:-1]
    assert abs(o.compute_error(2) - np.sqrt(np.mean(r ** 2)) / np.sqrt(np.mean(fc ** 2))) < 1e-13
    # norm 1: numerator over the core, denominator over the whole padded array (poisson_d3.cc:437)
    assert abs(o.compute_error(1) - np.max(np.abs(r)) / np.max(np.abs(R))) < 1e-13
    assert o.compute_error(0) == 0.0


# ------------------------------------------------------------------------------------------------
# singular problems: null vectors and projection
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("domain,mesh", [("PPN", "UUD"), ("PPP", "UUU"), ("NNN", "UUD")])
def test_left_null_vector_annihilates_operator(domain, mesh):
    pb = Problem(ind=(2, 2, 3), domain=domain, mesh=mesh, depth=1)
    o = pb.oracle()
    A = _dense_operator(o)
    n = pb.n
    ws = []
    for d in range(3):
        ax = po.grid_axis(n[d], 1.0, pb.gtype[d], 1.0)
        ws.append(po.null_vector_1d(n[d], pb.per[d], 1.0 / (n[d] - 1), ax["xix2"][1:-1], ax["xixx"][1:-1]))
    w = np.einsum("i,j,k->ijk", *ws).ravel()
    assert np.max(np.abs(w @ A)) < 1e-9 * np.max(np.abs(A))
    assert np.max(np.abs(A @ np.ones(A.shape[1]))) < 1e-9 * np.max(np.abs(A))


def test_projected_solve_reaches_tolerance():
    pb = Problem(ind=(5, 5, 5), domain="PPN", mesh="UUD", semantics=po.SEM_PROJECTED, rel_tol=1e-10)
    for sm in (po.SMOOTH_RB, po.SMOOTH_LEX):
        o = pb.oracle(smoother=sm)
        p, hist = o.solve(np.zeros(pb.shape), pb.random_rhs(2))
        assert hist[-1] < 1e-10 and len(hist) < 60

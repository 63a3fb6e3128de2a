"""Independent pins of three oracle functions, written as functions of the oracle library so that
tests/test_oracle_mutations.py can run the same checks against deliberately mutated builds.
Each pin raises AssertionError when the oracle departs from the property it checks.

  * tseries' Nusselt number (lib/io/tseries.cc:195-246): closed forms with flow -- an x cosine
    mode (duplicated periodic node included), the z-face pair (w linear in z), the collocated
    z metric read at node indices on a tanh grid (pad value included), both kappa choices of
    scalar.cc:69-86.
  * mgSolve in the reference's semantics (lib/poisson/poisson.cc:100-151): the mean removed over
    the *padded* residual array after updatePads, the stop on |r_(k-1) - r_k| < mgTolerance with
    r_0 = 0 read from the history, resType 0 giving one cycle.
  * the residual pads of padMode 1 (mirror at Neumann walls, wrap for periodic z; not a reference
    rule): full weighting reproduces cosine modes at every coarse node, boundary rows included,
    and on Neumann boundaries R = W_c^-1 P^T W_f / 8 with trapezoid weights W.
"""
import numpy as np

from oracle import pyoracle as po
from helpers import Problem


# ------------------------------------------------------------------------------------------------
# Nusselt number
# ------------------------------------------------------------------------------------------------
def _kappa(rbc_type, Ra, Pr):
    # scalar.cc:69-86
    return {1: 1.0, 2: 1.0 / np.sqrt(Pr * Ra), 3: 1.0 / Pr, 4: 1.0 / np.sqrt(Pr * Ra)}[rbc_type]


def _rbc(library, ind, L, mesh, rbc_type=2, Ra=1e7, Pr=1.0, beta=(1.0, 1.0, 1.0)):
    p = po.ts_params(ind=ind, L=L, domain="PPN", mesh=mesh, beta=beta, scalar=True, prob_type=5,
                     rbc_type=rbc_type, Ra=Ra, Pr=Pr, Re=1.0, semantics=po.SEM_PROJECTED)
    return po.OracleTS(p, library=library)


def nusselt_xmode(library=None):
    """w = W cos(2 pi x / Lx) on every z face, T = cos(2 pi x / Lx), uniform x: the x sum of
    cos^2 over the N nodes 0..N-1 (the periodic node N-1 duplicating node 0) is (N-1)/2 + 1, every
    node has the same x width, so Nu = 1 + W ((N-1)/2 + 1) / (N kappa) for any y and z grid."""
    for rbc_type, Ra, Pr in ((1, 2e3, 0.7), (2, 1e7, 0.7)):
        o = _rbc(library, (4, 3, 4), (2.0, 1.0, 1.0), "UUD", rbc_type, Ra, Pr, beta=(1.0, 1.0, 1.4))
        Vz, T = o.array(po.VZ), o.array(po.TEMP)
        N = T.shape[0] - 2
        th = 2.0 * np.pi * np.arange(-1, N + 1) / (N - 1)
        W = 0.37
        Vz[...] = (W * np.cos(th))[:, None, None]
        T[...] = np.cos(th)[:, None, None]
        want = 1.0 + W * ((N - 1) / 2.0 + 1.0) / N / _kappa(rbc_type, Ra, Pr)
        got = o.nusselt()
        o.close()
        assert abs(got - want) <= 1e-12 * abs(want), (rbc_type, got, want)


def nusselt_face_pair(library=None):
    """w on the z faces = the face coordinate z_k = (k + 1/2) h (pads included), T = 1, uniform
    grid: the node value (w_(k-1) + w_k) / 2 is the node coordinate k h, whose volume mean over
    the N equal-width nodes is Lz / 2, so Nu = 1 + (Lz / 2) / kappa."""
    Lz = 1.5
    o = _rbc(library, (3, 3, 4), (1.0, 1.0, Lz), "UUU", 1, 2e3, 0.7)
    Vz, T = o.array(po.VZ), o.array(po.TEMP)
    N = T.shape[2] - 2
    h = Lz / (N - 1)
    Vz[...] = ((np.arange(-1, N) + 0.5) * h)[None, None, :]
    T[...] = 1.0
    want = 1.0 + (Lz / 2.0) / _kappa(1, 2e3, 0.7)
    got = o.nusselt()
    o.close()
    assert abs(got - want) <= 1e-12 * abs(want), (got, want)


def nusselt_tanh_metric(library=None):
    """w = 1, T = 1 at node plane k0 only, tanh z grid: Nu - 1 = dz_k0 / sum_k dz_k / kappa with
    dz_k = dZt (dz/dzeta) at the face coordinate zeta = (k + 1/2) dZt -- the collocated metric read
    at node index k (tseries.cc:205-237) -- and, at k = N-1 beyond the last face, the uniform value
    Lz dZt that the tanh grid leaves in the pad (grid.cc:442-519)."""
    Lz, beta = 1.0, 1.3
    for k0 in (0, 5, 16):
        o = _rbc(library, (3, 3, 4), (1.0, 1.0, Lz), "UUD", 2, 1e7, 1.0, beta=(1.0, 1.0, beta))
        Vz, T = o.array(po.VZ), o.array(po.TEMP)
        N = T.shape[2] - 2
        dzt = 1.0 / (N - 1)
        zc = (np.arange(N) + 0.5) * dzt
        dz = dzt * Lz * beta / np.cosh(beta * (1.0 - 2.0 * zc)) ** 2 / np.tanh(beta)
        dz[N - 1] = Lz * dzt
        Vz[...] = 1.0
        T[...] = 0.0
        T[:, :, k0 + 1] = 1.0
        want = 1.0 + dz[k0] / dz.sum() / _kappa(2, 1e7, 1.0)
        got = o.nusselt()
        o.close()
        assert abs(got - want) <= 1e-12 * abs(want), (k0, got, want)


# ------------------------------------------------------------------------------------------------
# mgSolve, reference semantics
# ------------------------------------------------------------------------------------------------
def _wrap_xy(a):
    """updatePads on one rank with periodic x and y: full planes, x first (the exchange order of
    poisson_d3.cc:671-686), so the y exchange carries the x pads into the corners."""
    a[0, :, :], a[-1, :, :] = a[-3, :, :], a[2, :, :]
    a[:, 0, :], a[:, -1, :] = a[:, -3, :], a[:, 2, :]
    return a


def reference_mean_over_padded(library=None):
    """After loading (vcCount 0: no cycle runs), residualData(0) holds the right-hand side with its
    x/y pads exchanged, z pads 0, minus the mean over the whole padded array (poisson.cc:114-129)."""
    pb = Problem(ind=(3, 3, 4), domain="PPN", mesh="UUD", semantics=po.SEM_REFERENCE, vc_count=0, res_type=1)
    o = pb.oracle(library=library)
    rng = np.random.default_rng(11)
    f = rng.standard_normal(pb.shape) + 0.5
    o.solve(np.zeros(pb.shape), f)
    want = np.zeros(pb.shape)
    want[1:-1, 1:-1, 1:-1] = f[1:-1, 1:-1, 1:-1]
    want = _wrap_xy(want)
    # the padded mean and the core mean differ here by far more than the tolerance
    assert abs(want.mean() - want[1:-1, 1:-1, 1:-1].mean()) > 1e-3
    want -= want.mean()
    got = o.array(po.R, 0)
    assert np.max(np.abs(got - want)) < 1e-14, np.max(np.abs(got - want))
    o.close()


def reference_stop_rule(library=None):
    """The cycle count is the first k with |r_(k-1) - r_k| < mgTolerance, r_0 = 0 (poisson.cc:102,
    144-146), read from the history; the history's last value is computeError(1) of the returned
    solution (max |f' - L p| over the core / max |f'| over the padded array, f' the loaded RHS)."""
    for tol in (1e-2, 1e-4):
        pb = Problem(ind=(4, 4, 4), domain="PPN", mesh="UUD", semantics=po.SEM_REFERENCE, vc_count=40,
                     res_type=1, mg_tol=tol)
        o = pb.oracle(library=library)
        f = pb.random_rhs(4)
        p, hist = o.solve(np.zeros(pb.shape), f)
        o.close()
        n = len(hist)
        assert 1 <= n < 40
        prev = np.concatenate([[0.0], hist[:-1]])
        d = np.abs(prev - hist)
        assert d[-1] < tol and np.all(d[:-1] >= tol), (tol, hist)
        # computeError(1) of the returned iterate, from an independent evaluation of the residual
        pb.vc_count = 0
        q = pb.oracle(library=library)
        q.solve(np.zeros(pb.shape), f)          # no cycle: loads f' = padded f minus its padded mean
        fp = q.array(po.R, 0).copy()
        q.close()
        r = pb.oracle(library=library)
        r.level = 0
        r.array(po.P, 0)[...] = p
        r.array(po.R, 0)[...] = fp
        r.compute_residual()
        rmax = float(np.max(np.abs(r.array(po.T, 0)[1:-1, 1:-1, 1:-1])))
        r.close()
        assert abs(rmax / np.max(np.abs(fp)) - hist[-1]) <= 1e-10 * hist[-1]


def reference_restype0_one_cycle(library=None):
    """resType 0: computeError returns 0, |0 - 0| < mgTolerance after the first cycle."""
    pb = Problem(ind=(3, 3, 3), domain="PPN", mesh="UUD", semantics=po.SEM_REFERENCE, vc_count=20, res_type=0)
    o = pb.oracle(library=library)
    p, hist = o.solve(np.zeros(pb.shape), pb.random_rhs(1))
    o.close()
    assert len(hist) == 1 and hist[0] == 0.0


# ------------------------------------------------------------------------------------------------
# padMode 1 residual pads
# ------------------------------------------------------------------------------------------------
def _restrict_via_residual(o, f):
    """tmp = f - L 0 with the pads of computeResidual, then coarsen: the coarse right-hand side."""
    o.level = 0
    o.array(po.P, 0)[...] = 0.0
    o.array(po.R, 0)[...] = f
    o.compute_residual()
    o.coarsen()
    return o.array(po.R, 1).copy()


def pad_mode1_cosine_modes(library=None):
    """Full weighting is the tensor product of (1/4, 1/2, 1/4), so it maps g = prod_d cos(theta_d i_d)
    to g(2K) prod_d (1 + cos theta_d) / 2 wherever the pads hold g's own values: Neumann modes
    theta = pi m / M are symmetric about each wall (g(-1) = g(1)), periodic modes theta = 2 pi m / M
    wrap (g(-1) = g(M-1)).  With padMode 1 this holds at every coarse node, boundary rows included;
    the reference's zero pads (padMode 0) miss it on the non-periodic-x/y boundary rows."""
    for domain in ("NNN", "PPP", "PPN"):
        for pad_mode in (1, 0):
            pb = Problem(ind=(4, 3, 4), domain=domain, mesh="UUU", depth=1, pad_mode=pad_mode)
            o = pb.oracle(library=library)
            n = pb.n
            idx = [np.arange(-1, k + 1) for k in n]
            ths = [(2.0 if per else 1.0) * np.pi * m / (k - 1) for per, m, k in zip(pb.per, (2, 1, 3), n)]
            g = np.einsum("i,j,k->ijk", *[np.cos(t * i) for t, i in zip(ths, idx)])
            got = _restrict_via_residual(o, g)[1:-1, 1:-1, 1:-1]
            o.close()
            cidx = [np.arange((k - 1) // 2 + 1) for k in n]
            want = np.einsum("i,j,k->ijk", *[np.cos(t * 2 * i) for t, i in zip(ths, cidx)])
            want *= np.prod([(1.0 + np.cos(t)) / 2.0 for t in ths])
            err = np.max(np.abs(got - want))
            if pad_mode == 1:
                assert err < 1e-14, (domain, err)
            else:
                # zero pads: exact inside, wrong on the boundary rows of a non-x/y-periodic direction
                inner = got[1:-1, 1:-1, 1:-1] - want[1:-1, 1:-1, 1:-1]
                assert np.max(np.abs(inner)) < 1e-14
                assert err > 1e-3, (domain, err)


def pad_mode1_weighted_transpose(library=None):
    """Neumann in every direction, padMode 1: the restriction matrix R (coarse x fine, through
    computeResidual's pads) equals W_c^-1 P^T W_f / 8 on every row, P the trilinear prolongation
    matrix and W the trapezoid weights (1/2 per boundary index), i.e. full weighting is the
    transpose of prolongation in the trapezoid inner product.  Both matrices by brute force."""
    pb = Problem(ind=(3, 2, 3), domain="NNN", mesh="UUU", depth=1, pad_mode=1)
    o = pb.oracle(library=library)
    nf, nc = o.dims(0), o.dims(1)
    R = np.zeros((int(np.prod(nc)), int(np.prod(nf))))
    for q, idx in enumerate(np.ndindex(*nf)):
        e = np.zeros(pb.shape)
        e[tuple(i + 1 for i in idx)] = 1.0
        R[:, q] = _restrict_via_residual(o, e)[1:-1, 1:-1, 1:-1].ravel()
    Pm = np.zeros((int(np.prod(nf)), int(np.prod(nc))))
    for q, idx in enumerate(np.ndindex(*nc)):
        o.level = 1
        c = o.array(po.P, 1)
        c[...] = 0.0
        c[tuple(i + 1 for i in idx)] = 1.0
        o.prolong()
        Pm[:, q] = o.array(po.P, 0)[1:-1, 1:-1, 1:-1].ravel()
    o.close()

    def trap(shape):
        ws = []
        for k in shape:
            w = np.ones(k)
            w[0] = w[-1] = 0.5
            ws.append(w)
        return np.einsum("i,j,k->ijk", *ws).ravel()

    want = (Pm.T * trap(nf)[None, :]) / trap(nc)[:, None] / 8.0
    assert np.max(np.abs(R - want)) < 1e-15, np.max(np.abs(R - want))


# ------------------------------------------------------------------------------------------------
# field operators on caller arrays (or_fd_*): closed forms on polynomials in the computational
# coordinates.  Node coordinates are xi = i h, face coordinates xi = (i + 1/2) h (index -1..n);
# central differences are exact on quadratics, two- and four-point face averages on linear data.
# ------------------------------------------------------------------------------------------------
_FD_GRID = dict(ind=(4, 3, 4), L=(1.0, 2.0, 1.0), domain="NPN", mesh="DUD", beta=(1.1, 1.0, 1.3))


def _fd_oracle(library):
    return po.OracleTS(po.ts_params(**_FD_GRID), library=library)


def _fd_coords(o, loc):
    """computational coordinates of every padded entry of location loc (0 node, 1..3 faces)"""
    sh = o.loc_shape(loc)
    h = [1.0 / (1 << n) for n in (o.params.xInd, o.params.yInd, o.params.zInd)]
    ax = []
    for d in range(3):
        face = loc == d + 1
        ax.append(h[d] * (np.arange(sh[d]) - 1 + (0.5 if face else 0.0)))
    return np.meshgrid(*ax, indexing="ij")


def _fd_metric(o, loc, d, q):
    node = loc == 0 or loc - 1 != d
    return o.metric(d, node, q)[:o.loc_shape(loc)[d]].reshape([-1 if a == d else 1 for a in range(3)])


def fd_diff_quadratic(library=None):
    """computeDiff of F = sum_d c_d xi_d^2 adds sum_d (xi_dd 2 c_d xi_d + 0.5 xi_d^2 2 c_d) to the
    core of H and leaves H's pads as they were (a node field and the x faces)."""
    o = _fd_oracle(library)
    rng = np.random.default_rng(31)
    for loc in (0, 1):
        X = _fd_coords(o, loc)
        c = (1.0, -2.0, 0.5)
        F = sum(c[d] * X[d] ** 2 for d in range(3))
        H0 = rng.standard_normal(F.shape)
        H = H0.copy()
        o.fd_diff(loc, F, H)
        add = sum(_fd_metric(o, loc, d, 2) * 2 * c[d] * X[d] + 0.5 * _fd_metric(o, loc, d, 3) * 2 * c[d]
                  for d in range(3)) * np.ones(F.shape)
        exp = H0.copy()
        exp[1:-1, 1:-1, 1:-1] += add[1:-1, 1:-1, 1:-1]
        assert np.allclose(H, exp, rtol=1e-11, atol=1e-10 * np.max(np.abs(add))), loc
    o.close()


def fd_nlin_linear(library=None):
    """computeNLin of F = sum_d a_d xi_d advected by V_d = U_d + b_d xi_d + g_d xi_e (e the next
    direction) subtracts sum_d Vbar_d xi_d,x a_d, Vbar_d the same linear function at the location
    of F: the interpolation slices are averages of face values around that location."""
    o = _fd_oracle(library)
    a, U, b, g = (0.7, -1.3, 2.1), (1.5, -0.25, 0.8), (0.3, -0.6, 0.45), (-0.2, 0.35, 0.15)
    V = []
    for d in range(3):
        Xd = _fd_coords(o, d + 1)
        V.append(U[d] + b[d] * Xd[d] + g[d] * Xd[(d + 1) % 3])
    for loc in (0, 1, 2, 3):
        X = _fd_coords(o, loc)
        F = sum(a[d] * X[d] for d in range(3))
        H = np.zeros(F.shape)
        o.fd_nlin(loc, F, V, H)
        sub = sum((U[d] + b[d] * X[d] + g[d] * X[(d + 1) % 3]) * _fd_metric(o, loc, d, 1) * a[d] for d in range(3))
        exp = np.zeros(F.shape)
        exp[1:-1, 1:-1, 1:-1] = -sub[1:-1, 1:-1, 1:-1]
        assert np.allclose(H, exp, rtol=1e-11, atol=1e-12 * np.max(np.abs(sub))), loc
    o.close()


def fd_gradient_divergence_linear(library=None):
    """gradient of F = sum_d a_d xi_d is xi_c(face) a_c on each face core and leaves the rest; the
    divergence of V_c = b_c xi_c is sum_c xi_c(node) b_c on the node core and 0 on its pads."""
    o = _fd_oracle(library)
    X = _fd_coords(o, 0)
    a = (0.3, -0.9, 1.7)
    F = sum(a[d] * X[d] for d in range(3))
    G = [np.full(o.loc_shape(c + 1), 7.0) for c in range(3)]
    o.fd_gradient(F, G)
    for c in range(3):
        exp = np.full(G[c].shape, 7.0)
        val = _fd_metric(o, c + 1, c, 1) * a[c] * np.ones(G[c].shape)
        exp[1:-1, 1:-1, 1:-1] = val[1:-1, 1:-1, 1:-1]
        assert np.allclose(G[c], exp, rtol=1e-12, atol=1e-13), c
    b = (1.1, 0.4, -2.2)
    V = [b[c] * _fd_coords(o, c + 1)[c] for c in range(3)]
    div = o.fd_divergence(V)
    val = sum(_fd_metric(o, 0, c, 1) * b[c] for c in range(3)) * np.ones(div.shape)
    exp = np.zeros(div.shape)
    exp[1:-1, 1:-1, 1:-1] = val[1:-1, 1:-1, 1:-1]
    assert np.allclose(div, exp, rtol=1e-12, atol=1e-13)
    o.close()


def fd_bc_rules(library=None):
    """Wall rules on Vx of the NPN grid: x walls (faces, FC) Dirichlet 0.5 at x- (mean of image and
    first face), Neumann at x+ (image = last face); y periodic (syncData: ghost = the core layer one
    in from the far side, the periodic node duplicated); z walls (nodes, CC) Dirichlet 1.5 at z- (the
    end node set), Neumann at z+ (end node = its neighbour)."""
    o = _fd_oracle(library)
    rng = np.random.default_rng(5)
    F = rng.standard_normal(o.loc_shape(1))
    o.fd_bc(1, [1, 2, 0, 0, 1, 2], [0.5, 0.0, 0.0, 0.0, 1.5, 0.0], F)
    ny = F.shape[1] - 2
    assert np.allclose((F[0, 2:-2, 2:-2] + F[1, 2:-2, 2:-2]) / 2, 0.5, atol=1e-14)
    assert np.array_equal(F[-1, 2:-2, 2:-2], F[-2, 2:-2, 2:-2])
    assert np.array_equal(F[2:-2, 0, 2:-2], F[2:-2, ny - 1, 2:-2])      # F(-1) = F(M - 1), M = ny - 1
    assert np.array_equal(F[2:-2, -1, 2:-2], F[2:-2, 2, 2:-2])          # F(M + 1) = F(1)
    assert np.all(F[:, :, 1] == 1.5)
    assert np.array_equal(F[:, :, -2], F[:, :, -3])
    o.close()

"""Shared problem builders: the same inputs for the oracle (CPU checker) and the HIP path."""
import numpy as np

from oracle import pyoracle as po


def axes_oracle(n, L, gtype, beta):
    return [po.grid_axis(n[d], L[d], gtype[d], beta[d]) for d in range(3)]


def metrics_from_axes(ax):
    return {"xix2": ax[0]["xix2"], "xixx": ax[0]["xixx"], "ety2": ax[1]["xix2"],
            "etyy": ax[1]["xixx"], "ztz2": ax[2]["xix2"], "ztzz": ax[2]["xixx"]}


class Problem:
    """One multigrid configuration: index sizes, domain/mesh strings, solver settings."""

    def __init__(self, ind=(5, 5, 5), domain="PPN", mesh="UUD", L=(1.0, 1.0, 1.0), beta=(1.0, 1.0, 1.0),
                 depth=None, pre=2, post=2, vc_count=60, res_type=2, gs=1, test_poisson=False,
                 semantics=po.SEM_PROJECTED, mg_tol=1e-6, rel_tol=1e-10, pad_mode=0, coarse_project=0,
                 walls=None):
        self.ind = tuple(ind)
        self.n = [(1 << i) + 1 for i in ind]
        self.per = tuple(c == "P" for c in domain)
        self.gtype = tuple(1 if c == "U" else 2 for c in mesh)
        self.L, self.beta = tuple(L), tuple(beta)
        self.depth = depth if depth is not None else min(ind) - 1
        self.pre, self.post, self.vc_count, self.res_type, self.gs = pre, post, vc_count, res_type, gs
        self.test_poisson, self.semantics = test_poisson, semantics
        self.mg_tol, self.rel_tol, self.pad_mode, self.coarse_project = mg_tol, rel_tol, pad_mode, coarse_project
        self.walls = walls
        self.shape = tuple(k + 2 for k in self.n)

    # ---- oracle ----
    def oracle(self, smoother=po.SMOOTH_RB, library=None):
        ax = axes_oracle(self.n, self.L, self.gtype, self.beta)
        p = po.Params()
        p.xInd, p.yInd, p.zInd = self.ind
        p.vcDepth, p.vcCount = self.depth, self.vc_count
        p.preSmooth, p.postSmooth, p.resType = self.pre, self.post, self.res_type
        if not self.gs:
            smoother = po.SMOOTH_JACOBI_RB if smoother == po.SMOOTH_RB else po.SMOOTH_JACOBI
        p.smoother = smoother
        p.xPer, p.yPer, p.zPer = (int(x) for x in self.per)
        p.testPoisson, p.semantics = int(self.test_poisson), self.semantics
        p.mgTolerance, p.relTolerance = self.mg_tol, self.rel_tol
        p.dXi, p.dEt, p.dZt = (1.0 / (k - 1) for k in self.n)
        p.padMode, p.coarseProject = self.pad_mode, self.coarse_project
        return po.OracleMG(p, metrics_from_axes(ax), walls=self.walls, library=library)

    # ---- HIP path ----
    def gpu(self, precision=0, coarse_direct=0, inner=(0, 0)):
        from saras_amd import mg
        pr = mg.Parser(xInd=self.ind[0], yInd=self.ind[1], zInd=self.ind[2], Lx=self.L[0], Ly=self.L[1],
                       Lz=self.L[2], xGrid=self.gtype[0], yGrid=self.gtype[1], zGrid=self.gtype[2],
                       betaX=self.beta[0], betaY=self.beta[1], betaZ=self.beta[2], xPer=self.per[0],
                       yPer=self.per[1], zPer=self.per[2], vcDepth=self.depth, vcCount=self.vc_count,
                       preSmooth=self.pre, postSmooth=self.post, gsSmooth=self.gs, resType=self.res_type,
                       mgTolerance=self.mg_tol, testPoisson=self.test_poisson, semantics=self.semantics,
                       padMode=self.pad_mode, coarseProject=self.coarse_project, relTolerance=self.rel_tol,
                       precision=precision, coarseDirect=coarse_direct, innerPreSmooth=inner[0],
                       innerPostSmooth=inner[1])
        return mg.Multigrid(mg.Grid(pr), pr, walls=self.walls)

    def random_rhs(self, seed=0):
        rng = np.random.default_rng(seed)
        f = np.zeros(self.shape)
        f[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(self.n))
        return f


def r2_walls(n, L=(1.0, 1.0, 1.0)):
    """Dirichlet planes of the TEST_POISSON case u = r^2/6 (poisson_d3.cc:513-572, single rank):
    the same plane on both sides of each direction, zero pads."""
    h = [1.0 / (k - 1) for k in n]
    M = [k - 1 for k in n]
    half = [M[0] // 2, M[1] // 2]
    xw = np.zeros((n[1] + 2, n[2] + 2))
    yw = np.zeros((n[0] + 2, n[2] + 2))
    zw = np.zeros((n[0] + 2, n[1] + 2))
    xd = h[0] + L[0] / 2.0
    for j in range(n[1]):
        yd = h[1] * (j - half[1])
        for k in range(n[2]):
            zd = h[2] * (k - M[2] // 2)
            xw[j + 1, k + 1] = (xd * xd + yd * yd + zd * zd) / 6.0
    yd = h[1] + L[1] / 2.0
    for i in range(n[0]):
        xd_ = h[0] * (i - half[0])
        for k in range(n[2]):
            zd = h[2] * (k - M[2] // 2)
            yw[i + 1, k + 1] = (xd_ * xd_ + yd * yd + zd * zd) / 6.0
    zd = h[2] + L[2] / 2.0
    for i in range(n[0]):
        xd_ = h[0] * (i - half[0])
        for j in range(n[1]):
            yd_ = h[1] * (j - half[1])
            zw[i + 1, j + 1] = (xd_ * xd_ + yd_ * yd_ + zd * zd) / 6.0
    return [xw, xw, yw, yw, zw, zw]


def r2_exact(n):
    """Analytic r^2/6 on the core nodes (poisson.cc:166-190)."""
    h = [1.0 / (k - 1) for k in n]
    M = [k - 1 for k in n]
    i = (np.arange(n[0]) - M[0] // 2) * h[0]
    j = (np.arange(n[1]) - M[1] // 2) * h[1]
    k = (np.arange(n[2]) - M[2] // 2) * h[2]
    X, Y, Z = np.meshgrid(i, j, k, indexing="ij")
    return (X * X + Y * Y + Z * Z) / 6.0


def sin_problem(n):
    """Config C1: f = -3 pi^2 sin(pi x) sin(pi y) sin(pi z) on the unit cube, u = 0 on the faces,
    Dirichlet data imposed at the ghost nodes from the analytic solution."""
    xs = [np.arange(-1, k + 1) / (k - 1) for k in n]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    u = np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)
    f = -3.0 * np.pi ** 2 * u
    f[0, :, :] = f[-1, :, :] = 0.0
    f[:, 0, :] = f[:, -1, :] = 0.0
    f[:, :, 0] = f[:, :, -1] = 0.0
    walls = []
    for d in range(3):
        for side in (0, -1):
            sl = [slice(None)] * 3
            sl[d] = side
            w = u[tuple(sl)].copy()
            # zero pads of the wall plane (the reference allocates them as 0)
            w[0, :] = w[-1, :] = 0.0
            w[:, 0] = w[:, -1] = 0.0
            walls.append(np.ascontiguousarray(w))
    return f, u, walls


def rel_close(a, b, rtol, scale=None):
    """max |a - b| <= rtol * scale (scale defaults to max |b|)."""
    s = float(np.max(np.abs(b))) if scale is None else float(scale)
    err = float(np.max(np.abs(a - b)))
    return err <= rtol * max(s, 1e-300), err, s

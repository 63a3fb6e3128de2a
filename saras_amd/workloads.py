"""Synthetic inputs of the benchmark and parity workloads.

All random draws come from a self-contained SplitMix64 stream (uniforms from the top 53 bits,
normals by Box-Muller), so any implementation of the generator reproduces the same fields.
"""
import math

import numpy as np

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & _M64

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & _M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self):
        """(0, 1): top 53 bits, offset by half an ulp so 0 never occurs."""
        return ((self.next_u64() >> 11) + 0.5) * (1.0 / (1 << 53))

    def randint(self, lo, hi):
        """Integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))

    def normal(self):
        u1, u2 = self.uniform(), self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def fourier_modes(seed, count=8, kmax=16):
    """Mode list (a, kx, ky, kz, phx, phy) of the stretched-grid Poisson workload."""
    g = SplitMix64(seed)
    modes = []
    for _ in range(count):
        kx, ky, kz = g.randint(1, kmax), g.randint(1, kmax), g.randint(1, kmax)
        a = g.normal()
        phx, phy = 2.0 * math.pi * g.uniform(), 2.0 * math.pi * g.uniform()
        modes.append((a, kx, ky, kz, phx, phy))
    return modes


def fourier_rhs(x, y, z, seed, count=8, kmax=16):
    """Padded RHS on node coordinates x, y, z (padded 1D arrays): a sum of `count` modes
    a cos(2 pi kx x + phx) cos(2 pi ky y + phy) cos(pi kz z), periodic in x, y over [0, 1] and
    Neumann-compatible in z; zero on the ghost layer."""
    f = np.zeros((len(x), len(y), len(z)))
    xc, yc, zc = x[1:-1], y[1:-1], z[1:-1]
    core = np.zeros((len(xc), len(yc), len(zc)))
    for a, kx, ky, kz, phx, phy in fourier_modes(seed, count, kmax):
        cx = np.cos(2.0 * math.pi * kx * xc + phx)
        cy = np.cos(2.0 * math.pi * ky * yc + phy)
        cz = np.cos(math.pi * kz * zc)
        core += a * cx[:, None, None] * cy[None, :, None] * cz[None, None, :]
    f[1:-1, 1:-1, 1:-1] = core
    return f


def stretched_poisson(ind=8, beta=1.0, seed=0, depth=None):
    """The 257^3 stretched-grid Poisson workload (index 8): unit cube, periodic in x and y,
    Neumann in z, tanh z grid with parameter beta, V(2,2) red-black cycles to a relative residual
    of 1e-10 with the right-hand side projected onto the operator's range.  Returns
    (Parser, Grid, rhs)."""
    from . import mg
    p = mg.Parser.from_strings(mesh="UUD", domain="PPN", xInd=ind, yInd=ind, zInd=ind, betaZ=beta,
                               vcDepth=(ind - 1) if depth is None else depth, vcCount=200, preSmooth=2,
                               postSmooth=2, gsSmooth=1, resType=2, mgTolerance=1e-6,
                               semantics=mg._L.SR_SEM_PROJECTED, relTolerance=1e-10)
    g = mg.Grid(p)
    f = fourier_rhs(g.xStaggr, g.yStaggr, g.zStaggr, seed)
    return p, g, f


class EvolvingPoisson:
    """A sequence of right-hand sides on the stretched-grid workload, as a time-stepping run
    produces them: the seeded modes of `fourier_rhs` with amplitudes a_m cos(w_m t + psi_m),
    w_m uniform in [0.5, 2], psi_m uniform in [0, 2 pi), at t = step * dt.  The solutions form a
    smooth sequence in t, so the previous solutions carry information about the next one."""

    def __init__(self, ind=8, beta=1.0, seed=0, dt=0.05, depth=None):
        self.p, self.g, _ = stretched_poisson(ind=ind, beta=beta, seed=seed, depth=depth)
        self.dt = dt
        modes = fourier_modes(seed)
        gen = SplitMix64(seed + 7919)
        self.omega = [0.5 + 1.5 * gen.uniform() for _ in modes]
        self.psi = [2.0 * math.pi * gen.uniform() for _ in modes]
        g = self.g
        xc, yc, zc = g.xStaggr[1:-1], g.yStaggr[1:-1], g.zStaggr[1:-1]
        self.shape = (len(g.xStaggr), len(g.yStaggr), len(g.zStaggr))
        self.fields = np.empty((len(modes), len(xc) * len(yc) * len(zc)))
        self.amp = []
        for m, (a, kx, ky, kz, phx, phy) in enumerate(modes):
            cx = np.cos(2.0 * math.pi * kx * xc + phx)
            cy = np.cos(2.0 * math.pi * ky * yc + phy)
            cz = np.cos(math.pi * kz * zc)
            self.fields[m] = (cx[:, None, None] * cy[None, :, None] * cz[None, None, :]).ravel()
            self.amp.append(a)

    def coefficients(self, step):
        t = step * self.dt
        return np.array([a * math.cos(w * t + s) for a, w, s in zip(self.amp, self.omega, self.psi)])

    def rhs(self, step):
        f = np.zeros(self.shape)
        f[1:-1, 1:-1, 1:-1] = (self.coefficients(step) @ self.fields).reshape(
            tuple(k - 2 for k in self.shape))
        return f

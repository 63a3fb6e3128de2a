"""Pins of the scalar_d3 part of the time-step oracle (temperature equation, buoyancy, RBC walls).
CPU only.

  * Crank-Nicolson diffusion of a sine mode on the conductive profile: T = 1 - z + e sin(pi z)
    evolves as 1 - z + e g^n sin(pi z) with the discrete growth factor g = (1 - a) / (1 + a),
    a = dt kappa lambda / 2, lambda = (2 - 2 cos(pi h)) / h^2 (no flow: the buoyancy of a
    horizontally uniform T is a gradient and the projection removes it)
  * buoyantForce::addForcing (buoyantForce.cc:59-68) and the advection of T by a uniform V, closed forms
  * scalar_d3::solveT's fixed point with the collocated z metric read at nodes (scalar_d3.cc:541)
  * the probType 5 temperature walls of scalar::initTBC (scalar.cc:196-228)"""
import numpy as np
import pytest

from oracle import pyoracle as po


def rbc(ind=(4, 4, 5), L=(1.0, 1.0, 1.0), domain="PPN", mesh="UUU", **kw):
    base = dict(scalar=True, prob_type=5, rbc_type=1, Ra=1e3, Pr=1.0, dt=2e-4, Re=1.0,
                semantics=po.SEM_PROJECTED, rel_tol=1e-12, vc_count=60, pre=2, post=2, res_type=2)
    base.update(kw)
    return po.OracleTS(po.ts_params(ind=ind, L=L, domain=domain, mesh=mesh, **base))


def test_crank_nicolson_diffusion_of_a_sine_mode():
    o = rbc(cn_tol=1e-13)
    T = o.array(po.TEMP)
    n = T.shape[2] - 2
    z = o.metric(2, True, 0)[: T.shape[2]]
    eps = 0.1
    for c in range(3):
        o.array(c)[...] = 0.0
    T[...] = (1.0 - z + eps * np.sin(np.pi * z))[None, None, :]
    o.array(po.PRES)[...] = 0.0
    o.impose_tbc()
    h = 1.0 / (n - 1)
    lam = (2.0 - 2.0 * np.cos(np.pi * h)) / h ** 2
    a = 2e-4 * 1.0 * lam / 2.0
    g = (1.0 - a) / (1.0 + a)
    steps = 10
    for _ in range(steps):
        assert o.time_advance() is None
    want = (1.0 - z + eps * g ** steps * np.sin(np.pi * z))[1:-1]
    got = o.array(po.TEMP)[1:-1, 1:-1, 1:-1]
    assert np.max(np.abs(got - want[None, None, :])) < 1e-9
    # no flow develops: the buoyancy of a horizontally uniform temperature is a pure gradient
    assert max(np.max(np.abs(o.array(c))) for c in range(3)) < 1e-8
    # conduction only: Nu - 1 = <w T> / kappa vanishes
    assert abs(o.nusselt() - 1.0) < 1e-8
    o.close()


def test_buoyancy_and_temperature_advection_closed_forms():
    o = rbc(ind=(4, 4, 4), domain="PPP", dt=1e-3)
    rng = np.random.default_rng(0)
    T = o.array(po.TEMP)
    T[...] = rng.standard_normal(T.shape)
    for c in range(3):
        o.array(c)[...] = 0.0
    o.array(po.PRES)[...] = 0.0
    o.build_rhs()
    Hz = o.array(po.HZ)
    # Vz faces k sit between nodes k and k+1; Fb = Ra Pr for rbcType 1
    want = 1e-3 * 1e3 * (T[1:-1, 1:-1, 1:-2] + T[1:-1, 1:-1, 2:-1]) / 2.0
    assert np.allclose(Hz[1:-1, 1:-1, 1:-1], want, rtol=1e-14, atol=1e-15)
    # advection of T = x by a uniform Vx = U: dT/dt = -U on a uniform grid (d2 of a linear field = 0)
    U = 0.7
    o.array(po.VX)[...] = U
    x = o.metric(0, True, 0)[: T.shape[0]]
    T[...] = x[:, None, None]
    o.build_trhs()
    R = o.array(po.TRHS)
    assert np.allclose(R[1:-1, 1:-1, 1:-1], (x[:, None, None] - 1e-3 * U * np.ones(T.shape))[1:-1, 1:-1, 1:-1],
                       rtol=1e-13, atol=1e-13)
    o.close()


def test_solve_t_fixed_point_uses_collocated_z_metric():
    o = rbc(ind=(4, 4, 5), mesh="UUD", beta=(1.0, 1.0, 1.3), dt=5e-3, cn_tol=1e-12)
    o.init_rbc(seed=3, amp=0.05)
    o.build_trhs()
    it = o.solve_t()
    assert it > 1
    T = o.array(po.TEMP)
    R = o.array(po.TRHS)
    n = T.shape
    hs = [1.0 / (k - 3) for k in n]
    mx = o.metric(0, True, 3)[1:-1]
    my = o.metric(1, True, 3)[1:-1]
    mz = o.metric(2, False, 3)[1: n[2] - 1]        # collocated array read at node indices
    c = T[1:-1, 1:-1, 1:-1]
    lap = (mx[:, None, None] * (T[2:, 1:-1, 1:-1] - 2 * c + T[:-2, 1:-1, 1:-1]) / hs[0] ** 2 +
           my[None, :, None] * (T[1:-1, 2:, 1:-1] - 2 * c + T[1:-1, :-2, 1:-1]) / hs[1] ** 2 +
           mz[None, None, :] * (T[1:-1, 1:-1, 2:] - 2 * c + T[1:-1, 1:-1, :-2]) / hs[2] ** 2)
    kappa = 1.0
    res = np.abs(c - 0.5 * 5e-3 * kappa * lap - R[1:-1, 1:-1, 1:-1])[:, :, 1:-1]    # bulk: z walls excluded
    assert res.max() < 1e-12
    # with the nodal z metric instead the residual is far from the tolerance on a stretched grid
    mzn = o.metric(2, True, 3)[1:-1]
    lapn = lap + (mzn - mz)[None, None, :] * (T[1:-1, 1:-1, 2:] - 2 * c + T[1:-1, 1:-1, :-2]) / hs[2] ** 2
    resn = np.abs(c - 0.5 * 5e-3 * kappa * lapn - R[1:-1, 1:-1, 1:-1])[:, :, 1:-1]
    assert resn.max() > 1e3 * res.max()
    o.close()


@pytest.mark.parametrize("domain", ["PPN", "NNN"])
def test_temperature_walls_probtype5(domain):
    o = rbc(ind=(4, 4, 4), domain=domain)
    T = o.array(po.TEMP)
    T[...] = np.random.default_rng(1).standard_normal(T.shape)
    o.impose_tbc()
    assert np.all(T[:, :, 1] == 1.0) and np.all(T[:, :, -2] == 0.0)   # hot bottom, cold top at the wall nodes
    if domain == "NNN":
        # adiabatic sides: the node outside the bulk copies its inner neighbour (interior of the wall)
        assert np.array_equal(T[1, 2:-2, 2:-2], T[2, 2:-2, 2:-2])
        assert np.array_equal(T[-2, 2:-2, 2:-2], T[-3, 2:-2, 2:-2])
        assert np.array_equal(T[2:-2, 1, 2:-2], T[2:-2, 2, 2:-2])
    else:
        # periodic sides: the pads copy the core across (periodicCC of the node grid)
        assert np.array_equal(T[0, 1:-1, 2:-2], T[-3, 1:-1, 2:-2])
        assert np.array_equal(T[-1, 1:-1, 2:-2], T[2, 1:-1, 2:-2])
    o.close()
